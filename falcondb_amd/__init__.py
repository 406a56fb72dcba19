"""falcondb_amd — MI355X (gfx950) execution backend for FalconDB's WHERE filter,
hash GROUP BY aggregation and hash join.

Native pieces (built by `python -m falcondb_amd.build`):
  libfalcon_gpu.so     the HIP kernels and the C ABI of include/falcon_gpu.h
  libfgpu_datagen.so   deterministic synthetic table generators

The Python layer mirrors the reference's operator names (see operators.py).
"""
from .operators import (AggExpr, AggFunc, BatchDictColumn, BatchGroupBy, ColumnRef, Context, DictColumn, Expr,
                        FalconError, GroupByExec, JoinIndex, JoinType, NullableColumn, RecordBatch, Table,
                        col, cmp, count_star, execute_agg_only, execute_hash_join,
                        execute_hash_join_projected, execute_hash_semi_anti_join, execute_select, execute_sort,
                        host_exact_sum, join_pairs,
                        lit_cmp, pack_validity)

__all__ = [
    "AggExpr", "AggFunc", "BatchDictColumn", "BatchGroupBy", "ColumnRef", "Context", "DictColumn", "Expr", "FalconError",
    "GroupByExec", "JoinIndex", "JoinType", "NullableColumn", "RecordBatch", "Table", "col",
    "cmp", "count_star", "execute_agg_only", "execute_hash_join", "execute_hash_join_projected",
    "execute_hash_semi_anti_join",
    "execute_select", "execute_sort", "host_exact_sum", "join_pairs", "lit_cmp", "pack_validity",
]
