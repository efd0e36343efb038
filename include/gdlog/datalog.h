// gdlog/datalog.h -- a Datalog front end: lowers a program over binary
// relations to LIEs of RA operators, the shape of the reference's
// hand-written programs (test/tc.cu, test/sg.cu, test/cspa.cu, which encode
// test/souffle/{tc,sg,cspa}.dl by hand), and runs them on the current engine
// context.  An extension: the reference builds its RA-operator lists by hand.
//
// Accepted (a subset of Souffle's syntax):
//   .decl R(a: number, b: number)       unary and binary relations of
//   .decl S(a: number)                  number / unsigned ids below 2^32
//   .input R, S(IO=file, ...)           relations given to run() (a filename
//                                       attribute is kept for drivers; the
//                                       other attributes are ignored)
//   .output R   .printsize R            relations whose tuples run() keeps
//                                       (every derived relation when a
//                                       program names none)
//   R(1, 2).  S(0).                     facts: ground atoms, added to the
//                                       relation's starting tuples
//   H(x, y) :- A(x, z), B(z, y), x != y.      R(y) :- E(0, y).
//     1 to 4 body atoms over variables (_ for an unused column) and
//     constants (one per binary atom, next to a variable); every head
//     variable bound by the body; no variable twice in one atom; each atom
//     after the first shares exactly one variable with what is joined so
//     far, and at most two variables are live between two joins; at most one
//     comparison (= != < <= > >=) of a head variable with another or with an
//     id.  Line and block comments.
// Constants in heads, negation, aggregates, functors, other arities and other
// directives are rejected (GDL_ERR_UNSUPPORTED), syntax errors and unknown
// relations with GDL_ERR_INVALID_ARG; the message names the line.  A tuple of
// a unary relation is the key c0 << 32.
//
// Lowering.  Strata are the SCCs of the relation dependency graph, in
// dependency order.  A stratum's rules whose bodies read no relation of the
// stratum form its init LIE (all reads FULL; a one-atom rule is a
// RelationalCopy).  Its other rules form the recursive LIE: one semi-naive
// variant per body atom of the stratum (that atom's DELTA joined with the
// FULL of the others), each a chain of RelationalJoins through binary tmp
// relations, then a RelationalCopy from DELTA per one-atom rule, then one
// RelationalACopy per mirror.  The engine joins on column 0, so each relation
// is stored in the column order its stratum reads most (a linear rule then
// needs no second copy), and a read in the other order goes to a reversed
// copy: a mirror kept by ACopy inside the relation's own recursive LIE, or a
// static copy built from FULL once the relation is final.  A body atom
// with a constant, E(0, y), becomes a selection relation #s(y) :- #c(w),
// E(w, y) over a one-fact relation #c(0); such relations are never kept.
// A chain joins next the atom that leaves the fewest variables live.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "gdl_hip.h"

namespace gdlog {

class DatalogProgram {
  public:
    DatalogProgram();
    ~DatalogProgram();
    DatalogProgram(const DatalogProgram &) = delete;
    DatalogProgram &operator=(const DatalogProgram &) = delete;

    // Parses and plans the program; on failure returns false with *status
    // and *err set, and the program stays empty.
    bool compile(const std::string &source, gdl_status *status, std::string *err);
    // Same for a Slog compiler's RA plan (the s-expression of the reference's
    // test/souffle/cspa.slogc): the plan's joins and copies are read back
    // into rules over the relations' logical columns (index-maintenance
    // acopies dropped, DELTA / FULL variants merged), relations no rule
    // derives become .input, derived ones not marked tmp become .output, and
    // the rules are compiled as above.  Slog names become identifiers
    // (value-flow -> value_flow).  source() returns the rules read back.
    bool compile_slog(const std::string &plan, gdl_status *status, std::string *err);
    // Runs every stratum on the current context.  inputs: packed keys
    // ((c0 << 32) | c1; c0 << 32 for a unary relation), any order, duplicates
    // allowed, for each .input relation.  Throws gdlog::Error (GDL_ERR_INVALID_ARG: not compiled, a
    // missing or unknown input; engine errors as raised).
    void run(const std::map<std::string, std::vector<uint64_t>> &inputs);
    // Sorted, duplicate-free keys of a kept relation after run(); nullptr for
    // a relation that is not kept.
    const std::vector<uint64_t> *result(const std::string &relation) const;
    // Kept relations, in declaration order.
    const std::vector<std::string> &kept() const;
    // .input relations, in declaration order.
    const std::vector<std::string> &inputs() const;
    // Declared arity of a relation (1 or 2; 0 if not declared).
    int arity(const std::string &relation) const;
    // The filename attribute of a relation's .input directive ("" if none).
    const std::string &input_file(const std::string &relation) const;
    // The lowered plan: one line per step, LIE and operator.
    const std::string &plan() const;
    // The Datalog text compiled (for compile_slog, the rules read back).
    const std::string &source() const;
    // Iteration count of each LIE of the last run, in plan order.
    const std::vector<int> &iterations() const;
    // New tuples of the last run: sum of |delta| over LIEs, iterations and
    // stored relations (a reversed mirror's tuples are not counted again).
    uint64_t derived() const;

    struct Impl;

  private:
    std::unique_ptr<Impl> impl_;
};

}  // namespace gdlog
