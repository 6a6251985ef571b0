// Physical plan structures: the fields of server/thrift/calciteserver.thrift
// (RelNode, ExprType, LiteralType, AggType, CollationType, PlanResult), read
// from the JSON document with the same field names.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "json.hpp"

namespace sdb {

enum class RelNodeType { TABLE_SCAN, FILTER, JOIN, PROJECT, AGGREGATE, SORT, UNKNOWN };
enum class ExprOption { LITERAL, COLUMN, EXPR };
enum class LiteralOption { LITERAL, RANGE };

struct LiteralType {
    LiteralOption literalOption = LiteralOption::LITERAL;
    int64_t value = 0;
    std::vector<std::vector<std::string>> rangeSet;
};

struct ExprType {
    ExprOption exprType = ExprOption::LITERAL;
    int64_t input = 0;
    std::string name;
    std::string op;
    std::vector<ExprType> operands;
    LiteralType literal;
    std::string type;
};

struct AggType {
    std::string agg;
    std::vector<int64_t> operands;
    std::string name;
    std::string type;
    bool distinct = false;
};

struct CollationType {
    int64_t field = 0;
    std::string direction;
    std::string nulls;
};

struct RelNode {
    int64_t id = 0;
    RelNodeType relOp = RelNodeType::UNKNOWN;
    std::vector<std::string> tables;
    std::vector<int64_t> inputs;
    ExprType condition;
    std::string joinType;
    std::vector<std::string> fields;
    std::vector<ExprType> exprs;
    std::vector<int64_t> group;
    std::vector<AggType> aggs;
    std::vector<CollationType> collation;
};

struct PlanResult {
    std::vector<RelNode> rels;
};

PlanResult parse_plan(const std::string &json_text);

} // namespace sdb
