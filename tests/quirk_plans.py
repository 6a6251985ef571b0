"""Plans over the SSB tables that exercise the reference executor's quirks
(models/transient_table.hpp):

* the left fold of nested logical filters (845-860): a logical node passes
  its parent's op to its first operand and its own op to the others, so
  OR(a, AND(b, c)) is applied as ((a OR b) AND c);
* a FILTER placed after a join starts from parent op "" = NONE: its first
  comparison replaces the flags, dropping the join's result (845-860 with
  logical(NONE, a, b) = b, kernels/selection.hpp:91-92);
* LITERAL projections (1003-1024, a fill of every row) and the literal-first
  and literal-second arithmetic projections (1109-1203).

Plain helpers for tests; not a test module."""
from sycldb_amd import plans as PL
from sycldb_amd import schema as S

col, lit, expr = PL.col, PL.lit, PL.expr


def left_fold_or_and(qty=10, disc=2, price=5_000_000):
    """OR(discount = d, AND(quantity < q, extendedprice > p)); SUM(revenue)."""
    b = PL.PlanBuilder()
    b.scan("lineorder")
    b.filter(expr("OR", expr("=", col(S.LO_DISCOUNT), lit(disc)),
                  expr("AND", expr("<", col(S.LO_QUANTITY), lit(qty)),
                       expr(">", col(S.LO_EXTENDEDPRICE), lit(price)))))
    b.project([col(S.LO_REVENUE)], ["lo_revenue"])
    b.aggregate([], 0, "revenue")
    return b.plan()


def left_fold_and_or(qty=30, disc=(1, 9), price=9_000_000):
    """AND(quantity < q, OR(discount = d0, discount = d1, extendedprice > p))
    from a NONE parent: the first operand gets NONE, the OR node gets AND and
    passes it to its first operand, so this is applied as
    (((qty AND d0) OR d1) OR price); GROUP BY discount, SUM(quantity)."""
    b = PL.PlanBuilder()
    b.scan("lineorder")
    b.filter(expr("AND", expr("<", col(S.LO_QUANTITY), lit(qty)),
                  expr("OR", expr("=", col(S.LO_DISCOUNT), lit(disc[0])),
                       expr("=", col(S.LO_DISCOUNT), lit(disc[1])),
                       expr(">", col(S.LO_EXTENDEDPRICE), lit(price)))))
    b.project([col(S.LO_DISCOUNT), col(S.LO_QUANTITY)], ["lo_discount", "lo_quantity"])
    b.aggregate([0], 1, "qty")
    return b.plan()


def filter_after_join(year=1993, disc=2):
    """Semi join with date (one year), then FILTER discount = d: the filter's
    NONE parent overwrites the join's flags; SUM(extendedprice)."""
    b = PL.PlanBuilder()
    lo = b.scan("lineorder")
    b.scan("ddate")
    ddf = b.filter(expr("=", col(S.D_YEAR), lit(year)))
    b.join(lo, ddf, S.LO_ORDERDATE, S.D_DATEKEY, "semi")
    b.filter(expr("=", col(S.LO_DISCOUNT), lit(disc)))
    b.project([col(S.LO_EXTENDEDPRICE)], ["lo_extendedprice"])
    b.aggregate([], 0, "price")
    return b.plan()


def literal_first(qty=25):
    """FILTER quantity < q; PROJECT [7, quantity]; SUM(7) = 7 x the selected
    rows.  (parse_execution_info records no source for a LITERAL expression,
    operations/preprocessing.cpp:144-170, so the literal goes first: the
    aggregate's operand 0 then still names a recorded column.)"""
    b = PL.PlanBuilder()
    b.scan("lineorder")
    b.filter(expr("<", col(S.LO_QUANTITY), lit(qty)))
    b.project([lit(7), col(S.LO_QUANTITY)], ["seven", "lo_quantity"])
    b.aggregate([], 0, "s")
    return b.plan()


def literal_operands(qty=25, agg=1, group=None):
    """FILTER quantity < q; PROJECT [quantity * 3, 100 - discount,
    revenue / 1000, 7]; SUM of projected column `agg`, optionally grouped by
    projected column `group`.  A projected column's min/max are those of its
    column operand (models/models.hpp:797-833), so grouping by 100 - discount
    (values 90..100, recorded range 0..10) puts every key outside its range:
    its group slot is (key - 0) % 11, distinct for the 11 keys, and the
    groups come out in that slot order."""
    b = PL.PlanBuilder()
    b.scan("lineorder")
    b.filter(expr("<", col(S.LO_QUANTITY), lit(qty)))
    b.project([expr("*", col(S.LO_QUANTITY), lit(3), typ="INTEGER"),
               expr("-", lit(100), col(S.LO_DISCOUNT), typ="INTEGER"),
               expr("/", col(S.LO_REVENUE), lit(1000), typ="INTEGER"), lit(7)],
              ["qty3", "hundred_minus_disc", "rev_k", "seven"])
    b.aggregate([] if group is None else [group], agg, "s")
    return b.plan()


PLANS = {
    "left_fold_or_and": left_fold_or_and(),
    "left_fold_and_or": left_fold_and_or(),
    "filter_after_join": filter_after_join(),
    "literal_first": literal_first(),
    "col_times_lit": literal_operands(agg=0),
    "lit_minus_col": literal_operands(agg=1),
    "col_div_lit": literal_operands(agg=2),
    "group_by_lit_minus_col": literal_operands(agg=0, group=1),
}
