// JSON -> plan structures (the role of SycldbJsonConverter.java:10-163 and
// the Thrift deserialiser on the reference's C++ side).
#include "plan.hpp"

#include <stdexcept>

namespace sdb {
namespace {

using json::Value;

RelNodeType rel_type(std::string_view s)
{
    if (s == "TABLE_SCAN") return RelNodeType::TABLE_SCAN;
    if (s == "FILTER") return RelNodeType::FILTER;
    if (s == "JOIN") return RelNodeType::JOIN;
    if (s == "PROJECT") return RelNodeType::PROJECT;
    if (s == "AGGREGATE") return RelNodeType::AGGREGATE;
    if (s == "SORT") return RelNodeType::SORT;
    return RelNodeType::UNKNOWN;
}

std::vector<int64_t> int_list(const Value &v)
{
    std::vector<int64_t> out;
    for (const Value &x : v.as_arr()) out.push_back(x.as_int());
    return out;
}

ExprType expr_from(const Value &v)
{
    ExprType e;
    const std::string_view k = v.at("exprType").as_str();
    if (k == "COLUMN") {
        e.exprType = ExprOption::COLUMN;
        e.input = v.at("input").as_int();
        if (v.has("name") && v.at("name").kind == Value::String) e.name = std::string(v.at("name").s);
    } else if (k == "EXPR") {
        e.exprType = ExprOption::EXPR;
        e.op = std::string(v.at("op").as_str());
        if (v.has("operands")) {
            const Value::Range ops = v.at("operands").as_arr();
            e.operands.reserve(ops.size());
            for (const Value &o : ops) e.operands.push_back(expr_from(o));
        }
    } else if (k == "LITERAL") {
        e.exprType = ExprOption::LITERAL;
        const Value &l = v.at("literal");
        const std::string_view lo = l.has("literalOption") ? l.at("literalOption").as_str() : "LITERAL";
        if (lo == "RANGE") {
            e.literal.literalOption = LiteralOption::RANGE;
            for (const Value &r : l.at("rangeSet").as_arr()) {
                std::vector<std::string> row;
                for (const Value &x : r.as_arr()) row.push_back(x.kind == Value::Int ? std::to_string(x.i) : std::string(x.as_str()));
                e.literal.rangeSet.push_back(row);
            }
        } else {
            e.literal.literalOption = LiteralOption::LITERAL;
            e.literal.value = l.at("value").as_int();
        }
    } else {
        throw std::invalid_argument("plan: unknown exprType " + std::string(k));
    }
    if (v.has("type") && v.at("type").kind == Value::String) e.type = std::string(v.at("type").s);
    return e;
}

} // namespace

PlanResult parse_plan(const std::string &text)
{
    const json::Document doc(text);
    const Value &root = doc.root();
    PlanResult plan;
    const Value::Range rels = root.at("rels").as_arr();
    plan.rels.reserve(rels.size());
    for (const Value &r : rels) {
        RelNode rel;
        rel.id = r.at("id").as_int();
        rel.relOp = rel_type(r.at("relOp").as_str());
        if (r.has("tables"))
            for (const Value &t : r.at("tables").as_arr()) rel.tables.emplace_back(t.as_str());
        if (r.has("inputs")) rel.inputs = int_list(r.at("inputs"));
        if (r.has("condition") && r.at("condition").kind == Value::Object) rel.condition = expr_from(r.at("condition"));
        if (r.has("joinType") && r.at("joinType").kind == Value::String) rel.joinType = std::string(r.at("joinType").s);
        if (r.has("fields"))
            for (const Value &f : r.at("fields").as_arr()) rel.fields.emplace_back(f.kind == Value::String ? f.s : std::string_view());
        if (r.has("exprs")) {
            const Value::Range es = r.at("exprs").as_arr();
            rel.exprs.reserve(es.size());
            for (const Value &e : es) rel.exprs.push_back(expr_from(e));
        }
        if (r.has("group")) rel.group = int_list(r.at("group"));
        if (r.has("aggs")) {
            for (const Value &a : r.at("aggs").as_arr()) {
                AggType agg;
                if (a.has("agg")) {
                    const Value &n = a.at("agg");
                    agg.agg = std::string(n.kind == Value::Object ? n.at("name").as_str() : n.as_str());
                }
                if (a.has("operands")) agg.operands = int_list(a.at("operands"));
                if (a.has("name") && a.at("name").kind == Value::String) agg.name = std::string(a.at("name").s);
                if (a.has("type") && a.at("type").kind == Value::String) agg.type = std::string(a.at("type").s);
                if (a.has("distinct") && a.at("distinct").kind == Value::Bool) agg.distinct = a.at("distinct").b;
                rel.aggs.push_back(agg);
            }
        }
        if (r.has("collation")) {
            for (const Value &c : r.at("collation").as_arr()) {
                CollationType ct;
                ct.field = c.at("field").as_int();
                if (c.has("direction")) ct.direction = std::string(c.at("direction").as_str());
                if (c.has("nulls")) ct.nulls = std::string(c.at("nulls").as_str());
                rel.collation.push_back(ct);
            }
        }
        plan.rels.push_back(std::move(rel));
    }
    for (size_t i = 0; i < plan.rels.size(); i++)
        if (plan.rels[i].id != (int64_t)i) throw std::invalid_argument("plan: rel ids must be 0..n-1 in order");
    return plan;
}

} // namespace sdb
