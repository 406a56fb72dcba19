"""The SQL subset's parser (CPU): statements of the filter / GROUP BY / join /
ORDER BY ... LIMIT workloads parse into the plan shapes the planner expects;
statements outside the subset fail with the reference's error kinds."""
import pytest

from falcondb_amd import _abi
from falcondb_amd.operators import FalconError
from falcondb_amd.sql import ColName, parse


def test_groupby_statement():
    q = parse("SELECT dept, COUNT(*), SUM(score), AVG(salary) FROM t WHERE score > 50 GROUP BY dept")
    assert q.table == "t" and q.group_by == [ColName(None, "dept")]
    assert [(i.func, i.arg.name if i.arg else None) for i in q.items] == [
        (None, "dept"), (_abi.COUNT, None), (_abi.SUM, "score"), (_abi.AVG, "salary")]
    assert q.where == ("cmp", ColName(None, "score"), ">", 50, False)


def test_groupby_several_keys():
    q = parse("SELECT a, b, COUNT(*) FROM t GROUP BY a, t.b")
    assert q.group_by == [ColName(None, "a"), ColName("t", "b")]


def test_join_statement_with_aliases():
    q = parse("SELECT SUM(b.bpay) AS s, COUNT(*) FROM probe p LEFT OUTER JOIN build AS b ON p.pk = b.bk "
              "WHERE p.f < 30.5 AND (p.g = 'x''y' OR p.h IS NOT NULL);")
    assert q.alias == "p" and q.join[0] == _abi.LEFT and q.join[1:3] == ("build", "b")
    assert q.join[3] == ColName("p", "pk") and q.join[4] == ColName("b", "bk")
    assert q.items[0].alias == "s"
    assert q.where == ("and", ("cmp", ColName("p", "f"), "<", 30.5, False),
                       ("or", ("cmp", ColName("p", "g"), "=", "x'y", False), ("isnull", ColName("p", "h"), True)))


def test_exists_becomes_a_join_term():
    q = parse("SELECT SUM(probe.ppay), COUNT(*) FROM probe WHERE NOT EXISTS "
              "(SELECT 1 FROM build WHERE build.bk = probe.pk) AND probe.f >= -3")
    assert len(q.exists) == 1 and q.exists[0].negated and q.exists[0].table == "build"
    assert q.where == ("cmp", ColName("probe", "f"), ">=", -3, False)


def test_order_by_limit_and_literal_first():
    q = parse("select v2 from fact where 40 > f order by v2 desc limit 10")
    assert q.order_by == [(ColName(None, "v2"), True, True)] and q.limit == 10
    assert q.where == ("cmp", ColName(None, "f"), ">", 40, True)


def test_order_by_several_keys_nulls_offset():
    q = parse("SELECT a, b FROM t ORDER BY b DESC NULLS FIRST, a, c ASC NULLS LAST LIMIT 5 OFFSET 2")
    assert q.order_by == [(ColName(None, "b"), True, False), (ColName(None, "a"), False, False),
                          (ColName(None, "c"), False, True)]
    assert q.limit == 5 and q.offset == 2
    # first / last / nulls stay usable as column names
    q = parse("SELECT first, last FROM t ORDER BY last, nulls DESC")
    assert q.order_by == [(ColName(None, "last"), False, False), (ColName(None, "nulls"), True, True)]


@pytest.mark.parametrize("text,code", [
    ("SELECT a FROM t ORDER BY a NULLS MIDDLE", 4),
    ("SELECT a FROM t ORDER BY a LIMIT 3 OFFSET -1", 4),
    ("SELECT FROM t", 4),
    ("SELECT a FROM t WHERE a ~ 3", 4),
    ("SELECT SUM(*) FROM t", 4),
    ("SELECT a FROM t LIMIT 2.5", 4),
    ("SELECT a FROM t HAVING a > 1", 102),
    ("SELECT COUNT(*) FROM p JOIN b ON p.k < b.k", 102),
    ("SELECT COUNT(*) FROM p WHERE a = 1 OR EXISTS (SELECT 1 FROM b WHERE b.k = p.k)", 102),
])
def test_outside_the_subset(text, code):
    with pytest.raises(FalconError) as e:
        parse(text)
    assert e.value.code == code
