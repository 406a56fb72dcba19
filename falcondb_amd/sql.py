"""The SQL subset in front of the device operators: the statements of the
filter / GROUP BY / join / ORDER BY ... LIMIT workloads as a drop-in for
Database::sql (src/db.rs:242).

Plans follow the reference's step order (src/db.rs:1226-1744): FROM and JOIN,
WHERE on the FROM table (the reference's scan-time filter), GROUP BY, ORDER BY,
LIMIT, then projection in SELECT-list order.  The COUNT(*)-only query takes the
count fast path (src/db.rs:1439-1476), and a correlated [NOT] EXISTS becomes a
SEMI / ANTI join (src/query/sql.rs:586-897).  Output columns are named the
reference's way (expr_alias, src/query/sql.rs:780-788): a column by its name, a
qualified column by its last part, `AS` by the alias, anything else `col<i>`
for its position in the SELECT list.

Accepted:
    SELECT item, ... FROM t [alias] [[INNER | LEFT [OUTER]] JOIN b [alias] ON x.k = y.k]
        [WHERE pred] [GROUP BY col, ...] [ORDER BY col [ASC | DESC]] [LIMIT n]
    item := [t.]col [AS a] | COUNT(*) | AGG([t.]col) [AS a]      AGG: COUNT SUM AVG MIN MAX
    pred := [t.]col op literal | literal op [t.]col | [t.]col IS [NOT] NULL
            | pred AND pred | pred OR pred | (pred)
            | [NOT] EXISTS (SELECT 1 FROM b [alias] WHERE x.k = y.k)   (a top-level conjunct)
    op   := = | != | <> | < | <= | > | >=
Anything else raises FalconError (InvalidQuery, or Unsupported for shapes the
device path does not plan).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .operators import (AggExpr, Context, Expr, FalconError, GroupByExec, JoinIndex, RecordBatch, Table, cmp,
                        col, count_star, execute_agg_only, execute_hash_join,
                        execute_hash_join_projected, execute_select,
                        execute_sort, execute_sort_limit, lit_cmp)

PARSE, INVALID_QUERY, UNSUPPORTED = 4, 6, 102  # FalconError codes

_TOKEN = re.compile(r"""
    \s*(?:
      (?P<num>-?\d+\.\d*(?:[eE][-+]?\d+)?|-?\d+[eE][-+]?\d+|-?\.\d+(?:[eE][-+]?\d+)?|-?\d+)
    | (?P<str>'(?:[^']|'')*')
    | (?P<op><>|!=|<=|>=|=|<|>)
    | (?P<punct>[(),.*;])
    | (?P<name>[A-Za-z_][A-Za-z0-9_]*|"[^"]+")
    )""", re.X)

_KEYWORDS = {"SELECT", "FROM", "WHERE", "GROUP", "BY", "ORDER", "LIMIT", "ASC", "DESC", "AND", "OR",
             "NOT", "IS", "NULL", "JOIN", "INNER", "LEFT", "OUTER", "ON", "AS", "EXISTS", "LIKE", "ILIKE",
             "IN"}
_AGGS = {"COUNT": _abi.COUNT, "SUM": _abi.SUM, "AVG": _abi.AVG, "MIN": _abi.MIN, "MAX": _abi.MAX}


def _tokens(text: str) -> list[tuple[str, str]]:
    out, pos, text = [], 0, text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise FalconError(PARSE, f"SQL: cannot read '{text[pos:pos + 20]}'")
        pos = m.end()
        kind = m.lastgroup
        val = m.group(kind)
        if kind == "name" and not val.startswith('"') and val.upper() in _KEYWORDS:
            kind, val = "kw", val.upper()
        elif kind == "name" and val.startswith('"'):
            val = val[1:-1]
        out.append((kind, val))
    while out and out[-1] == ("punct", ";"):
        out.pop()
    return out


@dataclass
class ColName:
    table: str | None
    name: str


@dataclass
class Item:
    func: int | None          # None: a plain column
    arg: ColName | None       # None with func: COUNT(*)
    alias: str | None


@dataclass
class Exists:
    negated: bool
    table: str
    alias: str | None
    left: ColName
    right: ColName


@dataclass
class Query:
    items: list[Item]
    table: str
    alias: str | None
    join: tuple | None = None          # (type, table, alias, left ColName, right ColName)
    where: object = None               # predicate tree
    group_by: list | None = None  # ColName per key
    order_by: list | None = None       # [(ColName, desc, nulls_last)]
    limit: int | None = None
    offset: int = 0
    exists: list = field(default_factory=list)


class _Parser:
    def __init__(self, text: str):
        self.t = _tokens(text)
        self.i = 0

    def peek(self, k: int = 0):
        return self.t[self.i + k] if self.i + k < len(self.t) else ("end", "")

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def accept(self, kind: str, val: str | None = None) -> bool:
        k, v = self.peek()
        if k == kind and (val is None or v == val):
            self.i += 1
            return True
        return False

    def accept_word(self, word: str) -> bool:
        """Takes an unquoted name spelling `word` (a keyword only where it
        is read, so columns may still be called first, last ...)."""
        k, v = self.peek()
        if k == "name" and not v.startswith('"') and v.upper() == word:
            self.i += 1
            return True
        return False

    def expect(self, kind: str, val: str | None = None):
        k, v = self.take()
        if k != kind or (val is not None and v != val):
            raise FalconError(PARSE, f"SQL: expected {val or kind}, found '{v or 'end'}'")
        return v

    def colname(self) -> ColName:
        a = self.expect("name")
        if self.accept("punct", "."):
            return ColName(a, self.expect("name"))
        return ColName(None, a)

    def alias(self) -> str | None:
        if self.accept("kw", "AS"):
            return self.expect("name")
        if self.peek()[0] == "name":
            return self.take()[1]
        return None

    def item(self) -> Item:
        if self.accept("punct", "*"):  # SELECT *: every column (expanded by the planner)
            return Item(None, ColName(None, "*"), None)
        k, v = self.peek()
        if k == "name" and v.upper() in _AGGS and self.peek(1) == ("punct", "("):
            self.i += 2
            if self.accept("punct", "*"):
                if v.upper() != "COUNT":
                    raise FalconError(PARSE, f"SQL: {v}(*)")
                arg = None
            else:
                arg = self.colname()
            self.expect("punct", ")")
            return Item(_AGGS[v.upper()], arg, self.alias())
        return Item(None, self.colname(), self.alias())

    def literal(self):
        k, v = self.take()
        if k == "num":
            return float(v) if any(c in v for c in ".eE") else int(v)
        if k == "str":
            return v[1:-1].replace("''", "'")
        raise FalconError(PARSE, f"SQL: expected a literal, found '{v or 'end'}'")

    def operand_is_literal(self) -> bool:
        return self.peek()[0] in ("num", "str")

    def primary(self):
        if self.accept("punct", "("):
            e = self.disjunction()
            self.expect("punct", ")")
            return e
        negated = False
        if self.peek() == ("kw", "NOT") and self.peek(1) == ("kw", "EXISTS"):
            self.i += 1
            negated = True
        if self.accept("kw", "EXISTS"):
            self.expect("punct", "(")
            self.expect("kw", "SELECT")
            self.literal()
            self.expect("kw", "FROM")
            t = self.expect("name")
            a = self.alias() if self.peek() != ("kw", "WHERE") else None
            self.expect("kw", "WHERE")
            left = self.colname()
            if self.expect("op") != "=":
                raise FalconError(UNSUPPORTED, "SQL: EXISTS needs an equality correlation")
            right = self.colname()
            self.expect("punct", ")")
            return ("exists", Exists(negated, t, a, left, right))
        if self.operand_is_literal():
            lit = self.literal()
            op = self.expect("op")
            return ("cmp", self.colname(), op, lit, True)
        c = self.colname()
        if self.accept("kw", "IS"):
            neg = self.accept("kw", "NOT")
            self.expect("kw", "NULL")
            return ("isnull", c, neg)
        neg = False
        if self.peek() == ("kw", "NOT") and self.peek(1) in (("kw", "LIKE"), ("kw", "ILIKE"), ("kw", "IN")):
            self.i += 1
            neg = True
        if self.peek() in (("kw", "LIKE"), ("kw", "ILIKE")):
            ci = self.take()[1] == "ILIKE"
            pat = self.literal()
            if not isinstance(pat, str):
                raise FalconError(PARSE, "SQL: LIKE takes a string pattern")
            return ("like", c, pat, neg, ci)
        if self.accept("kw", "IN"):
            self.expect("punct", "(")
            lits = [self.literal()]
            while self.accept("punct", ","):
                lits.append(self.literal())
            self.expect("punct", ")")
            return ("in", c, lits, neg)
        if neg:
            raise FalconError(PARSE, "SQL: NOT takes LIKE, ILIKE or IN here")
        op = self.expect("op")
        return ("cmp", c, op, self.literal(), False)

    def conjunction(self):
        e = self.primary()
        while self.accept("kw", "AND"):
            e = ("and", e, self.primary())
        return e

    def disjunction(self):
        e = self.conjunction()
        while self.accept("kw", "OR"):
            e = ("or", e, self.conjunction())
        return e

    def query(self) -> Query:
        self.expect("kw", "SELECT")
        items = [self.item()]
        while self.accept("punct", ","):
            items.append(self.item())
        self.expect("kw", "FROM")
        table = self.expect("name")
        alias = self.alias() if self.peek()[0] == "name" or self.peek() == ("kw", "AS") else None
        q = Query(items, table, alias)
        jtype = None
        if self.accept("kw", "JOIN"):
            jtype = _abi.INNER
        elif self.accept("kw", "INNER"):
            self.expect("kw", "JOIN")
            jtype = _abi.INNER
        elif self.accept("kw", "LEFT"):
            self.accept("kw", "OUTER")
            self.expect("kw", "JOIN")
            jtype = _abi.LEFT
        if jtype is not None:
            bt = self.expect("name")
            ba = self.alias() if self.peek()[0] == "name" or self.peek() == ("kw", "AS") else None
            self.expect("kw", "ON")
            left = self.colname()
            if self.expect("op") != "=":
                raise FalconError(UNSUPPORTED, "SQL: joins on an equality only")
            q.join = (jtype, bt, ba, left, self.colname())
        if self.accept("kw", "WHERE"):
            q.where = self.disjunction()
        if self.accept("kw", "GROUP"):
            self.expect("kw", "BY")
            q.group_by = [self.colname()]
            while self.accept("punct", ","):
                q.group_by.append(self.colname())
        if self.accept("kw", "ORDER"):
            self.expect("kw", "BY")
            q.order_by = []
            while True:
                c = self.colname()
                desc = False
                if self.accept("kw", "DESC"):
                    desc = True
                else:
                    self.accept("kw", "ASC")
                # NULLS FIRST / LAST; the default puts NULLs last for DESC only
                # (translate_order_by, src/query/sql.rs:1480-1492)
                nulls_last = desc
                if self.accept_word("NULLS"):
                    if self.accept_word("FIRST"):
                        nulls_last = False
                    elif self.accept_word("LAST"):
                        nulls_last = True
                    else:
                        raise FalconError(PARSE, "SQL: NULLS takes FIRST or LAST")
                q.order_by.append((c, desc, nulls_last))
                if not self.accept("punct", ","):
                    break
        if self.accept("kw", "LIMIT"):
            n = self.literal()
            if not isinstance(n, int) or n < 0:
                raise FalconError(PARSE, "SQL: LIMIT takes a non-negative integer")
            q.limit = n
        if self.accept_word("OFFSET"):
            n = self.literal()
            if not isinstance(n, int) or n < 0:
                raise FalconError(PARSE, "SQL: OFFSET takes a non-negative integer")
            q.offset = n
        if self.peek()[0] != "end":
            raise FalconError(UNSUPPORTED, f"SQL: unsupported clause at '{self.peek()[1]}'")
        return q


def parse(text: str) -> Query:
    """Parses one statement of the subset; [NOT] EXISTS conjuncts of WHERE are
    moved to Query.exists (they plan as SEMI / ANTI joins)."""
    q = _Parser(text).query()

    def split(e):
        if e is None:
            return None
        if e[0] == "exists":
            q.exists.append(e[1])
            return None
        if e[0] == "and":
            a, b = split(e[1]), split(e[2])
            return b if a is None else a if b is None else ("and", a, b)
        if _has_exists(e):
            raise FalconError(UNSUPPORTED, "SQL: EXISTS only as a top-level AND term")
        return e
    q.where = split(q.where)
    if len(q.exists) > 1 or (q.exists and q.join):
        raise FalconError(UNSUPPORTED, "SQL: one join or one EXISTS per query")
    return q


def _has_exists(e) -> bool:
    if e[0] == "exists":
        return True
    if e[0] in ("and", "or"):
        return _has_exists(e[1]) or _has_exists(e[2])
    return False


def _output_name(item: Item, i: int) -> str:
    if item.alias:
        return item.alias
    if item.func is None:
        return item.arg.name
    return f"col{i}"


class Database:
    """Named device tables and the SQL entry point (Database::sql,
    src/db.rs:242).  Join indexes are built once per (build table, key)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.tables: dict[str, Table] = {}
        self._indexes: dict[tuple[str, str], JoinIndex] = {}

    def create_table(self, name: str, columns: dict, batch_rows: int = 65536) -> Table:
        t = Table(self.ctx, columns, batch_rows=batch_rows)
        self.register(name, t)
        return t

    def register(self, name: str, table: Table) -> None:
        self.tables[name.lower()] = table

    def drop(self, name: str) -> None:
        t = self.tables.pop(name.lower())
        for k in [k for k in self._indexes if k[0] == name.lower()]:
            self._indexes.pop(k).free()
        t.free()

    def table(self, name: str) -> Table:
        try:
            return self.tables[name.lower()]
        except KeyError:
            raise FalconError(INVALID_QUERY, f"table '{name}' not found") from None

    def index(self, name: str, key: str) -> JoinIndex:
        k = (name.lower(), key.lower())
        if k not in self._indexes:
            self._indexes[k] = JoinIndex(self.ctx, self.table(name), key)
        return self._indexes[k]

    # ── planning ──────────────────────────────────────────────────────────
    def sql(self, text: str) -> RecordBatch:
        q = parse(text)
        left = self.table(q.table)
        names_left = {q.table.lower()} | ({q.alias.lower()} if q.alias else set())
        right_name = right = None
        names_right: set[str] = set()
        jtype = None
        if q.join:
            jtype, right_name, ralias, jl, jr = q.join
        elif q.exists:
            ex = q.exists[0]
            jtype = _abi.ANTI if ex.negated else _abi.SEMI
            right_name, ralias, jl, jr = ex.table, ex.alias, ex.left, ex.right
        if right_name is not None:
            right = self.table(right_name)
            names_right = {right_name.lower()} | ({ralias.lower()} if ralias else set())

        def side(c: ColName) -> str:
            """'probe' or 'build' for a column reference."""
            if c.table is not None:
                t = c.table.lower()
                if t in names_left:
                    return "probe"
                if t in names_right:
                    return "build"
                raise FalconError(1, f"table or alias '{c.table}' not in the query")  # ColumnNotFound
            # A name both tables carry means the FROM table's column: the
            # join output keeps that one as <name> and calls the other
            # right.<name> (src/query/join.rs:877-886; tests/join_projection.rs:145-175).
            if _has_col(left, c.name):
                return "probe"
            in_right = right is not None and _has_col(right, c.name)
            if in_right:
                return "build"
            raise FalconError(1, f"column '{c.name}' not found")

        where = self._predicate(q.where, side) if q.where is not None else None
        star = any(it.func is None and it.arg.name == "*" for it in q.items)
        if star and right is None:
            q.items = _expand_star(q.items, left.names)
        names = [_output_name(it, i) for i, it in enumerate(q.items)]
        has_agg = any(it.func is not None for it in q.items)

        if right is not None:
            return self._plan_join(q, left, right, right_name, jtype, jl, jr, where, side, names)
        if q.group_by is not None or has_agg:
            return self._plan_aggregate(q, left, where, side, names)
        if q.order_by is not None:
            return self._plan_sort_limit(q, left, where, names)
        if q.limit is not None or q.offset:
            raise FalconError(UNSUPPORTED, "SQL: LIMIT / OFFSET without ORDER BY")
        return self._plan_select(q, left, where, names)

    def _predicate(self, e, side) -> Expr:
        return _to_expr(e, side)

    def _aggs(self, q: Query, side, names, skip_plain: bool = True) -> list[AggExpr]:
        out = []
        for it, nm in zip(q.items, names):
            if it.func is None:
                continue
            if it.arg is None:
                out.append(AggExpr(_abi.COUNT, None, nm))
            else:
                out.append(AggExpr(it.func, it.arg.name, nm, side=side(it.arg)))
        return out

    def _plan_aggregate(self, q, t, where, side, names) -> RecordBatch:
        if q.order_by is not None or q.limit is not None or q.offset:
            raise FalconError(UNSUPPORTED, "SQL: ORDER BY / LIMIT over an aggregate")
        aggs = self._aggs(q, side, names)
        plain = [(it, nm) for it, nm in zip(q.items, names) if it.func is None]
        if q.group_by is None:
            if plain:
                raise FalconError(INVALID_QUERY, f"column '{plain[0][0].arg.name}' must appear in GROUP BY")
            if len(aggs) == 1 and aggs[0].func == _abi.COUNT and aggs[0].input is None:
                n = count_star(self.ctx, t, where)  # the COUNT(*) fast path
                return RecordBatch([names[0]], ["Int64"], [np.array([n], np.int64)], [None], [_abi.COUNT])
            return execute_agg_only(self.ctx, t, aggs, where)
        keys = [c.name for c in q.group_by]
        low = [k.lower() for k in keys]
        for it, _ in plain:
            if it.arg.name.lower() not in low:
                raise FalconError(INVALID_QUERY, f"column '{it.arg.name}' must appear in GROUP BY")
        res = GroupByExec(keys, aggs).execute(self.ctx, t, where)
        # operator order: the keys, then the aggregates; SELECT order may differ
        cols, dts, valid, funcs, ai = [], [], [], [], len(keys)
        for it, nm in zip(q.items, names):
            j = low.index(it.arg.name.lower()) if it.func is None else ai
            if it.func is not None:
                ai += 1
            cols.append(res.columns[j])
            dts.append(res.dtypes[j])
            valid.append(res.validity[j])
            funcs.append(res.agg_funcs[j] if j < len(res.agg_funcs) else None)
        return RecordBatch(names, dts, cols, valid, funcs)

    def _plan_join(self, q, left, right, right_name, jtype, jl, jr, where, side, names) -> RecordBatch:
        sl, sr = side(jl), side(jr)
        if {sl, sr} != {"probe", "build"}:
            raise FalconError(INVALID_QUERY, "SQL: the join condition must relate the two tables")
        pk, bk = (jl.name, jr.name) if sl == "probe" else (jr.name, jl.name)
        semi = jtype in (_abi.SEMI, _abi.ANTI)
        index = self.index(right_name, bk)
        star = any(it.func is None and it.arg.name == "*" for it in q.items)
        if (q.group_by is None and q.order_by is None and q.limit is None and not q.offset and not star
                and all(it.func is not None for it in q.items)):
            # aggregates only: the join is fused with them, nothing is materialised
            aggs = self._aggs(q, side, names)
            if semi and any(a.side == "build" for a in aggs):
                raise FalconError(INVALID_QUERY, "SQL: EXISTS exposes no column of its subquery")
            return execute_hash_join(self.ctx, left, right, pk, bk, jtype, aggs, filter=where, index=index)
        # Otherwise the join's output is materialised with the columns the rest
        # of the plan reads (execute_hash_join_projected's keep set), and the
        # plan's remaining steps run over it.  A build column the probe table
        # also names is called right.<name> there (src/query/join.rs:877-886).
        def out_name(c: ColName) -> str:
            if side(c) == "probe":
                return left.names[left.column_index(c.name)]
            bname = right.names[right.column_index(c.name)]
            return "right." + bname if _has_col(left, bname) else bname
        refs = [it.arg for it in q.items if it.arg is not None and it.arg.name != "*"]
        refs += [c for c in (q.group_by or []) + [k[0] for k in q.order_by or []] if c is not None]
        probe_cols, build_cols = [], []
        for c in refs:
            lst, nm = (probe_cols, left.names[left.column_index(c.name)]) if side(c) == "probe" else \
                (build_cols, right.names[right.column_index(c.name)])
            if semi and lst is build_cols:
                raise FalconError(INVALID_QUERY, "SQL: EXISTS exposes no column of its subquery")
            if nm not in lst:
                lst.append(nm)
        if star:
            probe_cols = list(left.names)
            build_cols = [] if semi else list(right.names)
        jt = execute_hash_join_projected(self.ctx, left, right, pk, bk, jtype, probe_cols, build_cols,
                                         filter=where, index=index)
        try:
            def rename(c):
                return None if c is None else ColName(None, out_name(c))
            items = []
            for it in q.items:
                if it.func is None and it.arg.name == "*":
                    items += [Item(None, ColName(None, n), None) for n in jt.names]
                else:
                    items.append(Item(it.func, rename(it.arg), it.alias))
            q2 = Query(items, "", None, group_by=[rename(c) for c in q.group_by] if q.group_by else None,
                       order_by=[(rename(k[0]),) + tuple(k[1:]) for k in q.order_by] if q.order_by else None,
                       limit=q.limit, offset=q.offset)
            names2 = [_output_name(it, i) for i, it in enumerate(items)] if star else names

            def side2(c):
                return "probe"
            if q2.group_by is not None or any(it.func is not None for it in items):
                return self._plan_aggregate(q2, jt, None, side2, names2)
            if q2.order_by is not None:
                return self._plan_sort_limit(q2, jt, None, names2)
            return self._plan_select(q2, jt, None, names2)
        finally:
            jt.free()

    def _plan_sort_limit(self, q, t, where, names) -> RecordBatch:
        """ORDER BY k1, k2 ... [LIMIT n [OFFSET m]] (execute_sort_limit,
        src/query/sort.rs:59-115).  One key selected alone with LIMIT ≤ 32
        takes the fused top-k (fgpu_topk); every other shape sorts the
        passing rows on the device (fgpu_sort_limit) and gathers the selected
        columns in output order."""
        if any(it.func is not None for it in q.items):
            raise FalconError(INVALID_QUERY, "SQL: aggregates need GROUP BY or no plain columns")
        keys = [(k[0].name, k[1], k[2]) for k in q.order_by]
        lim = None if q.limit is None else q.limit + q.offset
        c, desc, nulls_last = keys[0]
        if (len(keys) == 1 and nulls_last == desc and lim is not None and lim <= 32 and len(q.items) == 1
                and q.items[0].arg.name.lower() == c.lower()
                and t.dtypes[t.column_index(c)] != _abi.UTF8):
            rows, vals = execute_sort_limit(self.ctx, t, c, lim, desc=desc, filter=where)
            dt = "Float64" if vals.dtype == np.float64 else "Int64"
            valid = None
            if t.column_stats(c)["nulls"]:
                # NULL rows follow the others for DESC and lead for ASC (the reference's placement)
                i = np.arange(vals.size)
                if desc:
                    nn = col(c).is_not_null()
                    valid = i < count_star(self.ctx, t, nn if where is None else where & nn)
                else:
                    isn = col(c).is_null()
                    valid = i >= min(vals.size, count_star(self.ctx, t, isn if where is None else where & isn))
            res = RecordBatch([names[0]], [dt], [vals], [valid], [None])
        else:
            _, res = execute_sort(self.ctx, t, keys, limit=lim, filter=where,
                                  columns=[it.arg.name for it in q.items])
            res = RecordBatch(names, res.dtypes, res.columns, res.validity, [None] * len(names))
        if q.offset:
            res = RecordBatch(res.names, res.dtypes, [a[q.offset:] for a in res.columns],
                              [None if v is None else v[q.offset:] for v in res.validity], res.agg_funcs)
        return res

    def _plan_select(self, q, t, where, names) -> RecordBatch:
        if any(it.func is not None for it in q.items):
            raise FalconError(INVALID_QUERY, "SQL: aggregates need GROUP BY or no plain columns")
        res = execute_select(self.ctx, t, where, [it.arg.name for it in q.items])
        return RecordBatch(names, res.dtypes, res.columns, res.validity, [None] * len(names))


def _to_expr(e, side=None) -> Expr:
    """A parsed WHERE tree as an Expr; with `side`, every column must belong
    to the FROM table (the reference filters it at scan time)."""
    kind = e[0]
    if kind == "and":
        return _to_expr(e[1], side) & _to_expr(e[2], side)
    if kind == "or":
        return _to_expr(e[1], side) | _to_expr(e[2], side)
    c = e[1]
    if side is not None and side(c) != "probe":
        raise FalconError(UNSUPPORTED, f"SQL: WHERE on '{c.name}' of the joined table")
    if kind == "isnull":
        return col(c.name).is_not_null() if e[2] else col(c.name).is_null()
    if kind == "like":
        return col(c.name).like(e[2], negated=e[3], case_insensitive=e[4])
    if kind == "in":
        return col(c.name).isin(e[2], negated=e[3])
    _, _, op, lit, lit_first = e
    op = "!=" if op == "<>" else op
    return lit_cmp(lit, op, c.name) if lit_first else cmp(c.name, op, lit)


def where_expr(text: str) -> Expr:
    """The Expr of a WHERE clause over one table, e.g.
    `where_expr("s LIKE 'b%' AND n NOT IN (1, 2)")`."""
    q = parse(f"SELECT COUNT(*) FROM t WHERE {text}")
    if q.exists:
        raise FalconError(UNSUPPORTED, "SQL: EXISTS needs a query")
    return _to_expr(q.where)


def _expand_star(items: list, columns: list[str]) -> list:
    out = []
    for it in items:
        if it.func is None and it.arg.name == "*":
            out += [Item(None, ColName(None, n), None) for n in columns]
        else:
            out.append(it)
    return out


def _has_col(t: Table, name: str) -> bool:
    try:
        t.column_index(name)
        return True
    except FalconError:
        return False
