// datalog.cpp -- the Datalog front end (include/gdlog/datalog.h) and its C
// ABI (include/gdl_datalog.h): parse, stratify, lower each stratum to an
// init LIE and a recursive LIE of RelationalJoin / RelationalCopy /
// RelationalACopy over packed binary relations, and run them.
#include "gdlog/datalog.h"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <functional>
#include <set>
#include <sstream>

#include "gdl_datalog.h"
#include "gdlog/exception.h"
#include "gdlog/lie.h"

namespace gdlog {

namespace {

// ---- lexer ------------------------------------------------------------------

struct Tok {
    enum Kind { Ident, Num, Str, Punct, End } kind;
    std::string s;
    int line;
};

struct CompileError {
    gdl_status status;
    std::string msg;
};

[[noreturn]] void fail(gdl_status st, int line, const std::string &msg) {
    throw CompileError{st, "line " + std::to_string(line) + ": " + msg};
}

std::vector<Tok> lex(const std::string &src) {
    std::vector<Tok> out;
    int line = 1;
    size_t i = 0;
    const size_t n = src.size();
    while (i < n) {
        const char c = src[i];
        if (c == '\n') {
            ++line;
            ++i;
        } else if (std::isspace((unsigned char)c)) {
            ++i;
        } else if (c == '/' && i + 1 < n && src[i + 1] == '/') {
            while (i < n && src[i] != '\n') ++i;
        } else if (c == '/' && i + 1 < n && src[i + 1] == '*') {
            const int start = line;
            i += 2;
            while (i + 1 < n && !(src[i] == '*' && src[i + 1] == '/')) {
                if (src[i] == '\n') ++line;
                ++i;
            }
            if (i + 1 >= n) fail(GDL_ERR_INVALID_ARG, start, "unterminated comment");
            i += 2;
        } else if (std::isalpha((unsigned char)c) || c == '_' || c == '?') {
            size_t j = i + 1;
            while (j < n && (std::isalnum((unsigned char)src[j]) || src[j] == '_' || src[j] == '?'))
                ++j;
            out.push_back({Tok::Ident, src.substr(i, j - i), line});
            i = j;
        } else if (std::isdigit((unsigned char)c)) {
            size_t j = i + 1;
            while (j < n && std::isalnum((unsigned char)src[j])) ++j;
            out.push_back({Tok::Num, src.substr(i, j - i), line});
            i = j;
        } else if (c == '"') {
            const int start = line;
            size_t j = i + 1;
            while (j < n && src[j] != '"') {
                if (src[j] == '\\' && j + 1 < n) ++j;
                if (src[j] == '\n') ++line;
                ++j;
            }
            if (j >= n) fail(GDL_ERR_INVALID_ARG, start, "unterminated string");
            out.push_back({Tok::Str, src.substr(i + 1, j - i - 1), start});
            i = j + 1;
        } else if (i + 1 < n && (src.compare(i, 2, ":-") == 0 || src.compare(i, 2, "!=") == 0 ||
                                 src.compare(i, 2, "<=") == 0 || src.compare(i, 2, ">=") == 0)) {
            out.push_back({Tok::Punct, src.substr(i, 2), line});
            i += 2;
        } else if (std::strchr(".(),:=<>!;[]{}+-*/%$@&|^~", c)) {
            out.push_back({Tok::Punct, std::string(1, c), line});
            ++i;
        } else {
            fail(GDL_ERR_INVALID_ARG, line, std::string("unexpected character '") + c + "'");
        }
    }
    out.push_back({Tok::End, "", line});
    return out;
}

// ---- program ----------------------------------------------------------------

struct Atom {
    int rel = -1;
    // variables; an anonymous one gets a name no other has; a unary atom
    // repeats its variable in v[1], so it joins and projects like a binary
    // one that is never read in reversed order
    std::string v[2];
};

struct Rule {
    int head = -1;
    int head_arity = 2;
    std::string hv[2];
    std::vector<Atom> body;
    int cmp_op = GDL_PRED_NONE;  // one comparison: cl <op> cr, or cl <op> cmp_value
    std::string cl, cr;
    bool cmp_const = false;
    uint32_t cmp_value = 0;
    int line = 0;
};

constexpr char kAnon[] = "_#";

bool is_anon(const std::string &v) { return v.compare(0, 2, kAnon) == 0; }

}  // namespace

// ---- plan -------------------------------------------------------------------

struct DatalogProgram::Impl {
    enum PhysKind { kBase, kMirror, kStaticRev, kTmp };
    struct Phys {
        std::string name;
        int rel = -1;      // logical relation (-1 for a tmp)
        bool rev = false;  // stored as (c1, c0)
        PhysKind kind = kBase;
    };
    struct Op {
        int kind = 0;  // 0 join, 1 copy, 2 acopy
        int src = -1, inner = -1, dst = -1;  // phys ids; src: join outer / copy source
        RelationVersion src_ver = FULL;
        gdl_colmap map{};
        gdl_pred pred{GDL_PRED_NONE, 0, 0};
    };
    struct Step {
        bool lie = false;
        int phys = -1;  // !lie: the reversed copy built from its relation's FULL
        std::string label;
        std::vector<int> updates, statics, tmps;
        std::vector<Op> ops;
    };

    // program
    std::vector<std::string> rel_names;
    std::map<std::string, int> rel_id;
    std::vector<int> arity;                   // 1 or 2
    std::vector<std::vector<uint64_t>> facts;  // ground atoms of the program, packed
    std::vector<char> is_input, is_kept;
    std::vector<char> hidden;  // made by the front end (constants), never kept
    std::map<int, std::string> input_file;  // an .input's filename attribute
    bool any_kept = false;
    std::vector<Rule> rules;
    // plan
    std::vector<Phys> phys;
    std::vector<int> base, other;
    std::vector<char> prim_rev, other_built;
    std::vector<Step> steps;
    std::string source;  // the Datalog text compiled
    std::string plan_text;
    std::vector<std::string> kept;          // relations run() keeps, declaration order
    std::vector<std::string> input_names;   // .input relations, declaration order
    bool compiled = false;
    // last run
    std::vector<int> iterations;
    uint64_t derived = 0;
    std::map<std::string, std::vector<uint64_t>> results;

    void parse(const std::vector<Tok> &toks);
    void lower();
    void describe();
    int phys_base(int rel);
    int new_phys(const std::string &name, int rel, bool rev, PhysKind kind);
    void run(const std::map<std::string, std::vector<uint64_t>> &inputs);
};

namespace {

// ---- parser -----------------------------------------------------------------

struct Parser {
    const std::vector<Tok> &t;
    DatalogProgram::Impl &m;
    size_t p = 0;
    int anon = 0;

    const Tok &peek(size_t k = 0) const { return t[std::min(p + k, t.size() - 1)]; }
    bool is(const char *s, size_t k = 0) const {
        const Tok &x = peek(k);
        return x.kind == Tok::Punct && x.s == s;
    }
    const Tok &next() {
        const Tok &x = peek();
        if (p + 1 < t.size()) ++p;
        return x;
    }
    std::string found() const {
        const Tok &x = peek();
        return x.kind == Tok::End ? " at the end of the program" : " before '" + x.s + "'";
    }
    void expect(const char *s) {
        if (!is(s)) fail(GDL_ERR_INVALID_ARG, peek().line, std::string("expected '") + s + "'" + found());
        next();
    }
    std::string ident(const char *what) {
        if (peek().kind != Tok::Ident)
            fail(GDL_ERR_INVALID_ARG, peek().line, std::string("expected ") + what + found());
        return next().s;
    }
    int relation(const std::string &name, int line) const {
        auto it = m.rel_id.find(name);
        if (it == m.rel_id.end()) fail(GDL_ERR_INVALID_ARG, line, "relation " + name + " is not declared");
        return it->second;
    }

    void program() {
        while (peek().kind != Tok::End) {
            if (is(".")) directive();
            else rule();
        }
    }

    void directive() {
        next();
        const int line = peek().line;
        const std::string d = ident("a directive");
        if (d == "decl") decl(line);
        else if (d == "input" || d == "output" || d == "printsize") io(d);
        else fail(GDL_ERR_UNSUPPORTED, line, "directive ." + d + " is not supported");
    }

    void decl(int line) {
        const std::string name = ident("a relation name");
        if (m.rel_id.count(name)) fail(GDL_ERR_INVALID_ARG, line, "relation " + name + " is declared twice");
        expect("(");
        int arity = 0;
        while (!is(")")) {
            ident("an attribute name");
            expect(":");
            const int tl = peek().line;
            const std::string ty = ident("an attribute type");
            if (ty != "number" && ty != "unsigned")
                fail(GDL_ERR_UNSUPPORTED, tl, "attribute type " + ty + " (number and unsigned only)");
            ++arity;
            if (!is(",")) break;
            next();
        }
        const int close_line = peek().line;
        expect(")");
        // a relation qualifier after the declaration, on its line
        static const char *quals[] = {"brie", "btree", "btree_delete", "eqrel", "choice",
                                      "inline", "no_inline", "magic", "no_magic", "override",
                                      "overridable", "nullary"};
        if (peek().kind == Tok::Ident && peek().line == close_line)
            for (const char *q : quals)
                if (peek().s == q)
                    fail(GDL_ERR_UNSUPPORTED, close_line,
                         "relation qualifier " + peek().s + " is not supported");
        if (arity != 1 && arity != 2)
            fail(GDL_ERR_UNSUPPORTED, line,
                 "relation " + name + " has arity " + std::to_string(arity) +
                     " (unary and binary relations only)");
        add_relation(name, arity, false);
    }

    int add_relation(const std::string &name, int arity, bool hidden) {
        m.rel_id[name] = (int)m.rel_names.size();
        m.rel_names.push_back(name);
        m.arity.push_back(arity);
        m.facts.emplace_back();
        m.is_input.push_back(0);
        m.is_kept.push_back(0);
        m.hidden.push_back(hidden);
        return (int)m.rel_names.size() - 1;
    }

    // (key = value, ...) after an I/O directive's relation; values are
    // identifiers, strings or numbers.  Only an .input's filename is used.
    std::map<std::string, std::string> attributes() {
        std::map<std::string, std::string> a;
        expect("(");
        while (!is(")")) {
            const std::string key = ident("an attribute name");
            expect("=");
            const Tok &v = peek();
            if (v.kind != Tok::Ident && v.kind != Tok::Str && v.kind != Tok::Num)
                fail(GDL_ERR_INVALID_ARG, v.line, "expected an attribute value" + found());
            a[key] = next().s;
            if (!is(",")) break;
            next();
        }
        expect(")");
        return a;
    }

    void io(const std::string &d) {
        for (;;) {
            const int line = peek().line;
            const int rel = relation(ident("a relation name"), line);
            if (d == "input") {
                m.is_input[rel] = 1;
            } else {
                m.is_kept[rel] = 1;
                m.any_kept = true;
            }
            if (is("(")) {
                const std::map<std::string, std::string> a = attributes();
                auto f = a.find("filename");
                if (d == "input" && f != a.end()) m.input_file[rel] = f->second;
            }
            if (!is(",")) break;
            next();
        }
    }

    uint32_t id_value(const Tok &x) {
        if (x.s.size() > 10 || x.s.find_first_not_of("0123456789") != std::string::npos ||
            std::stoull(x.s) > 0xffffffffull)
            fail(GDL_ERR_INVALID_ARG, x.line, "bad id " + x.s + " (decimal, below 2^32)");
        return (uint32_t)std::stoull(x.s);
    }

    std::string variable(const char *where) {
        const Tok &x = peek();
        if (x.kind == Tok::Num || x.kind == Tok::Str)
            fail(GDL_ERR_UNSUPPORTED, x.line, std::string("constants ") + where + " are not supported");
        if (x.kind != Tok::Ident) fail(GDL_ERR_INVALID_ARG, x.line, "expected a variable" + found());
        return next().s;
    }

    // An atom.  Its numbers go to consts (with their argument positions in
    // const_pos); without consts an atom may not hold numbers.
    Atom atom(std::vector<uint64_t> *consts = nullptr, std::vector<int> *const_pos = nullptr) {
        Atom a;
        const int line = peek().line;
        const std::string name = ident("a relation name");
        a.rel = relation(name, line);
        expect("(");
        int k = 0, nconst = 0;
        while (!is(")")) {
            const Tok &x = peek();
            if (x.kind == Tok::Num && consts) {
                consts->push_back(id_value(x));
                if (const_pos) const_pos->push_back(k);
                if (k < 2) a.v[k] = kAnon + std::to_string(++anon);
                next();
                ++nconst;
            } else {
                std::string v = variable("in rules");
                if (v == "_") v = kAnon + std::to_string(++anon);
                if (k < 2) a.v[k] = v;
            }
            ++k;
            if (peek().kind == Tok::Punct && std::strchr("+-*/%^&|", peek().s[0]))
                fail(GDL_ERR_UNSUPPORTED, peek().line, "arithmetic is not supported");
            if (!is(",")) break;
            next();
        }
        expect(")");
        const int want = m.arity[a.rel];
        if (k != want)
            fail(GDL_ERR_INVALID_ARG, line,
                 name + " takes " + std::to_string(want) + " argument" + (want > 1 ? "s" : "") +
                     ", not " + std::to_string(k));
        if (nconst && nconst != k && !const_pos)
            fail(GDL_ERR_UNSUPPORTED, line, "constants in rule heads are not supported");
        if (want == 1) a.v[1] = a.v[0];
        return a;
    }

    // A body atom R(.., c, ..) with one constant becomes #s(v), where
    //   #s(v) :- #c(w), R(.., w, ..).     #c(c).
    // a selection in its own rule, so the constant is matched by a join on
    // the constant's column before the atom meets the rest of the body.
    Atom select_constant(const Atom &a, const std::vector<uint64_t> &c,
                         const std::vector<int> &pos, int line) {
        if (m.arity[a.rel] == 1 || c.size() > 1)
            fail(GDL_ERR_UNSUPPORTED, line, "a ground atom in a rule body");
        const int p = pos[0], q = 1 - p;
        if (is_anon(a.v[q]))
            fail(GDL_ERR_UNSUPPORTED, line, "an atom holding a constant and no variable");
        const std::string cname = "#c" + std::to_string(c[0]);
        auto it = m.rel_id.find(cname);
        int crel;
        if (it == m.rel_id.end()) {
            crel = add_relation(cname, 1, true);
            m.facts[crel].push_back(c[0] << 32);
        } else {
            crel = it->second;
        }
        const std::string sname =
            "#s_" + m.rel_names[a.rel] + "_" + std::to_string(p) + "_" + std::to_string(c[0]);
        auto st = m.rel_id.find(sname);
        int srel;
        if (st == m.rel_id.end()) {
            srel = add_relation(sname, 1, true);
            Rule sel;
            sel.line = line;
            sel.head = srel;
            sel.head_arity = 1;
            sel.hv[0] = sel.hv[1] = "v";
            Atom src;
            src.rel = a.rel;
            src.v[q] = "v";
            src.v[p] = "w";
            Atom cw;
            cw.rel = crel;
            cw.v[0] = cw.v[1] = "w";
            sel.body = {cw, src};  // the one-tuple #c first: the outer of an init join
            m.rules.push_back(sel);
        } else {
            srel = st->second;
        }
        Atom out;
        out.rel = srel;
        out.v[0] = out.v[1] = a.v[q];
        return out;
    }

    void rule() {
        Rule r;
        r.line = peek().line;
        std::vector<uint64_t> ground;
        const Atom h = atom(&ground);
        if (!ground.empty()) {  // a fact
            if (!is("."))
                fail(GDL_ERR_UNSUPPORTED, r.line, "constants in rule heads are not supported");
            next();
            m.facts[h.rel].push_back((ground[0] << 32) | (ground.size() > 1 ? ground[1] : 0));
            return;
        }
        if (is(",")) fail(GDL_ERR_UNSUPPORTED, r.line, "rules with several heads are not supported");
        if (is(".")) fail(GDL_ERR_INVALID_ARG, r.line, "a fact with variables");
        expect(":-");
        for (;;) {
            const int line = peek().line;
            if (is("!")) fail(GDL_ERR_UNSUPPORTED, line, "negation is not supported");
            if (peek().kind == Tok::Ident && is("(", 1)) {
                std::vector<uint64_t> c;
                std::vector<int> pos;
                const Atom a = atom(&c, &pos);
                r.body.push_back(c.empty() ? a : select_constant(a, c, pos, line));
            } else {
                // a comparison of two variables, or of a variable and an id
                auto operand = [&](std::string *var, uint32_t *val) {
                    const Tok &x = peek();
                    if (x.kind == Tok::Num) {
                        *val = id_value(x);
                        next();
                        return false;
                    }
                    if (x.kind != Tok::Ident)
                        fail(GDL_ERR_INVALID_ARG, x.line, "expected a variable or an id" + found());
                    *var = next().s;
                    return true;
                };
                std::string lhs, rhs;
                uint32_t lc = 0, rc = 0;
                const bool lvar = operand(&lhs, &lc);
                static const std::pair<const char *, int> ops[] = {
                    {"=", GDL_PRED_EQ}, {"!=", GDL_PRED_NE}, {"<", GDL_PRED_LT},
                    {"<=", GDL_PRED_LE}, {">", GDL_PRED_GT}, {">=", GDL_PRED_GE}};
                int code = GDL_PRED_NONE;
                for (const auto &o : ops)
                    if (is(o.first)) code = o.second;
                if (code == GDL_PRED_NONE)
                    fail(GDL_ERR_INVALID_ARG, peek().line, "expected an atom or a comparison" + found());
                next();
                const bool rvar = operand(&rhs, &rc);
                if (r.cmp_op != GDL_PRED_NONE)
                    fail(GDL_ERR_UNSUPPORTED, line, "more than one comparison in a rule");
                if (!lvar && !rvar) fail(GDL_ERR_UNSUPPORTED, line, "a comparison of two ids");
                if (lhs == "_" || rhs == "_") fail(GDL_ERR_INVALID_ARG, line, "_ in a comparison");
                if (!lvar) {  // id <op> x  ==  x <flipped op> id
                    static const int flip[] = {GDL_PRED_NONE, GDL_PRED_EQ, GDL_PRED_NE, GDL_PRED_GT,
                                               GDL_PRED_GE, GDL_PRED_LT, GDL_PRED_LE};
                    code = flip[code];
                    r.cl = rhs;
                    r.cmp_const = true;
                    r.cmp_value = lc;
                } else if (!rvar) {
                    r.cl = lhs;
                    r.cmp_const = true;
                    r.cmp_value = rc;
                } else {
                    r.cl = lhs;
                    r.cr = rhs;
                }
                r.cmp_op = code;
            }
            if (is(";")) fail(GDL_ERR_UNSUPPORTED, peek().line, "disjunction is not supported");
            if (!is(",")) break;
            next();
        }
        expect(".");
        r.head = h.rel;
        r.head_arity = m.arity[h.rel];
        r.hv[0] = h.v[0];
        r.hv[1] = h.v[1];
        validate(r);
        m.rules.push_back(r);
    }

    void validate(const Rule &r) const {
        if (r.body.empty()) fail(GDL_ERR_UNSUPPORTED, r.line, "a rule needs a body atom");
        if (r.body.size() > 4) fail(GDL_ERR_UNSUPPORTED, r.line, "more than four body atoms");
        for (const Atom &a : r.body)
            if (m.arity[a.rel] == 2 && a.v[0] == a.v[1])
                fail(GDL_ERR_UNSUPPORTED, r.line, "variable " + a.v[0] + " twice in one atom");
        auto bound = [&](const std::string &v) {
            for (const Atom &a : r.body)
                if (a.v[0] == v || a.v[1] == v) return true;
            return false;
        };
        for (const std::string &v : r.hv) {
            if (is_anon(v)) fail(GDL_ERR_INVALID_ARG, r.line, "_ in a rule head");
            if (!bound(v))
                fail(GDL_ERR_INVALID_ARG, r.line, "head variable " + v + " is not bound by the body");
        }
        if (r.cmp_op != GDL_PRED_NONE) {
            for (const std::string &v : {r.cl, r.cmp_const ? r.cl : r.cr}) {
                if (!bound(v))
                    fail(GDL_ERR_INVALID_ARG, r.line, "variable " + v + " is not bound by the body");
                if (v != r.hv[0] && v != r.hv[1])
                    fail(GDL_ERR_UNSUPPORTED, r.line,
                         "comparison of " + v + ", which is not a head variable");
            }
        }
    }
};

// ---- lowering ---------------------------------------------------------------

// An operator over logical relations, before column orders are chosen.
struct AOp {
    int kind = 0;                 // 0 join, 1 copy
    int src_rel = -1;             // join outer / copy source (logical) ...
    bool src_rev = false;         // ... read in this column order (joins)
    int src_tmp = -1;             // ... or this tmp (join outer, NEWT)
    RelationVersion src_ver = FULL;
    int inner_rel = -1;           // join inner (FULL)
    bool inner_rev = false;
    int head = -1;                // output: the head relation, map in its logical order
    int dst_tmp = -1;             // ... or a tmp, map in its stored order
    gdl_colmap map{};
    gdl_pred pred{GDL_PRED_NONE, 0, 0};
};

// The rule's comparison over its head's columns.
gdl_pred rule_pred(const Rule &r) {
    if (r.cmp_op == GDL_PRED_NONE) return gdl_pred{GDL_PRED_NONE, 0, 0};
    const int l = r.cl == r.hv[0] ? 0 : 1;
    if (r.cmp_const) return gdl_pred{r.cmp_op | GDL_PRED_CONST, l, (int32_t)r.cmp_value};
    return gdl_pred{r.cmp_op, l, r.cr == r.hv[0] ? 0 : 1};
}

std::vector<std::string> shared_vars(const std::string cur[2], const Atom &a) {
    std::vector<std::string> s;
    for (int i = 0; i < 2; ++i) {
        if (i == 1 && cur[1] == cur[0]) break;
        if (a.v[0] == cur[i] || a.v[1] == cur[i]) s.push_back(cur[i]);
    }
    return s;
}

// Appends the joins that evaluate rule r starting from body atom s, read as
// version ver, then each next atom (in body order) that shares a variable
// with the tuples joined so far; intermediates go to tmps numbered from
// *ntmp, stored with the next join variable first.
void plan_chain(const Rule &r, int s, RelationVersion ver, std::vector<AOp> *ops, int *ntmp) {
    const int k = (int)r.body.size();
    std::vector<char> used(k, 0);
    used[s] = 1;
    std::string cur[2] = {r.body[s].v[0], r.body[s].v[1]};
    int cur_tmp = -1;
    // The next atom: one sharing a variable with c, the one that leaves the
    // fewest variables live after its join (a filter on bound variables
    // first), then the first in body order.
    auto pick = [&](const std::string c[2]) {
        int best = -1;
        size_t best_live = 0;
        for (int b = 0; b < k; ++b) {
            if (used[b] || shared_vars(c, r.body[b]).empty()) continue;
            std::set<std::string> need(r.hv, r.hv + 2);
            for (int x = 0; x < k; ++x)
                if (!used[x] && x != b) need.insert(r.body[x].v, r.body[x].v + 2);
            std::set<std::string> live;
            for (const std::string &v : {c[0], c[1], r.body[b].v[0], r.body[b].v[1]})
                if (need.count(v)) live.insert(v);
            if (best < 0 || live.size() < best_live) {
                best = b;
                best_live = live.size();
            }
        }
        return best;
    };
    auto join_var = [&](const std::string c[2], int b) {
        if (b < 0)
            fail(GDL_ERR_UNSUPPORTED, r.line, "body atoms that share no variable (a cross product)");
        const std::vector<std::string> sh = shared_vars(c, r.body[b]);
        if (sh.size() != 1)
            fail(GDL_ERR_UNSUPPORTED, r.line, "a join on two variables");
        return sh[0];
    };
    for (int step = 1; step < k; ++step) {
        const int b = pick(cur);
        const std::string j = join_var(cur, b);
        AOp op;
        std::string oo;  // the outer's other variable
        if (cur_tmp < 0) {
            op.src_rel = r.body[s].rel;
            op.src_rev = r.body[s].v[0] != j;
            op.src_ver = ver;
            oo = op.src_rev ? r.body[s].v[0] : r.body[s].v[1];
        } else {
            op.src_tmp = cur_tmp;  // stored with j first
            op.src_ver = NEWT;
            oo = cur[1];
        }
        const Atom &ia = r.body[b];
        op.inner_rel = ia.rel;
        op.inner_rev = ia.v[0] != j;
        const std::string io = op.inner_rev ? ia.v[0] : ia.v[1];
        used[b] = 1;
        auto src_of = [&](const std::string &v) {
            if (v == j) return (int)GDL_COL_OUTER0;
            if (v == oo) return (int)GDL_COL_OUTER1;
            if (v == io) return (int)GDL_COL_INNER1;
            fail(GDL_ERR_UNSUPPORTED, r.line, "variable " + v + " is not live at its join");
        };
        if (step == k - 1) {
            op.head = r.head;
            op.map = {r.head_arity, {src_of(r.hv[0]), src_of(r.hv[1])}};
            op.pred = rule_pred(r);
        } else {
            std::set<std::string> need(r.hv, r.hv + 2);
            for (int x = 0; x < k; ++x)
                if (!used[x]) need.insert(r.body[x].v, r.body[x].v + 2);
            std::vector<std::string> live;
            for (const std::string &v : {j, oo, io})
                if (need.count(v) && std::find(live.begin(), live.end(), v) == live.end())
                    live.push_back(v);
            if (live.size() > 2)
                fail(GDL_ERR_UNSUPPORTED, r.line, "three variables live between two joins");
            if (live.empty())
                fail(GDL_ERR_UNSUPPORTED, r.line, "body atoms that share no variable (a cross product)");
            std::string nxt[2] = {live[0], live.size() > 1 ? live[1] : live[0]};
            const std::string j2 = join_var(nxt, pick(nxt));
            if (nxt[0] != j2) std::swap(nxt[0], nxt[1]);
            op.dst_tmp = (*ntmp)++;
            op.map = {2, {src_of(nxt[0]), src_of(nxt[1])}};
            cur[0] = nxt[0];
            cur[1] = nxt[1];
            cur_tmp = op.dst_tmp;
        }
        ops->push_back(op);
    }
}

AOp plan_copy(const Rule &r, RelationVersion ver) {
    AOp op;
    op.kind = 1;
    op.pred = rule_pred(r);
    op.src_rel = r.body[0].rel;
    op.src_ver = ver;
    op.head = r.head;
    for (int i = 0; i < 2; ++i)
        op.map.src[i] = r.body[0].v[0] == r.hv[i] ? GDL_COL_INNER0 : GDL_COL_INNER1;
    op.map.arity = r.head_arity;
    return op;
}

// Strongly connected components of the head -> body graph (Tarjan), emitted
// dependencies first.
std::vector<std::vector<int>> components(int n, const std::vector<std::vector<int>> &adj) {
    std::vector<int> index(n, -1), low(n, 0), stack;
    std::vector<char> on(n, 0);
    std::vector<std::vector<int>> out;
    int counter = 0;
    std::function<void(int)> visit = [&](int v) {
        index[v] = low[v] = counter++;
        stack.push_back(v);
        on[v] = 1;
        for (int w : adj[v]) {
            if (index[w] < 0) {
                visit(w);
                low[v] = std::min(low[v], low[w]);
            } else if (on[w]) {
                low[v] = std::min(low[v], index[w]);
            }
        }
        if (low[v] == index[v]) {
            std::vector<int> comp;
            int w;
            do {
                w = stack.back();
                stack.pop_back();
                on[w] = 0;
                comp.push_back(w);
            } while (w != v);
            std::sort(comp.begin(), comp.end());
            out.push_back(comp);
        }
    };
    for (int v = 0; v < n; ++v)
        if (index[v] < 0) visit(v);
    return out;
}

gdl_colmap swap_outputs(gdl_colmap m) {
    std::swap(m.src[0], m.src[1]);
    return m;
}

int flip_source(int c) {
    return c == GDL_COL_INNER0 ? GDL_COL_INNER1 : c == GDL_COL_INNER1 ? GDL_COL_INNER0 : c;
}

const char *col_name(int c) {
    switch (c) {
    case GDL_COL_INNER0: return "i0";
    case GDL_COL_INNER1: return "i1";
    case GDL_COL_OUTER0: return "o0";
    default: return "o1";
    }
}

const char *ver_name(RelationVersion v) { return v == DELTA ? "DELTA" : v == FULL ? "FULL" : "NEWT"; }

const char *pred_name(int op) {
    static const char *n[] = {"", "=", "!=", "<", "<=", ">", ">=", "false"};
    return op >= 0 && op <= GDL_PRED_FALSE ? n[op] : "?";
}

}  // namespace

int DatalogProgram::Impl::new_phys(const std::string &name, int rel, bool rev, PhysKind kind) {
    phys.push_back({name, rel, rev, kind});
    return (int)phys.size() - 1;
}

int DatalogProgram::Impl::phys_base(int rel) {
    if (base[rel] < 0) {
        const char *order = arity[rel] == 1 ? "_1__1" : prim_rev[rel] ? "_2__2_1" : "_2__1_2";
        base[rel] = new_phys(rel_names[rel] + order, rel, prim_rev[rel], kBase);
    }
    return base[rel];
}

void DatalogProgram::Impl::parse(const std::vector<Tok> &toks) {
    Parser ps{toks, *this};
    ps.program();
}

void DatalogProgram::Impl::lower() {
    const int n = (int)rel_names.size();
    std::vector<std::vector<int>> adj(n);
    for (const Rule &r : rules)
        for (const Atom &a : r.body) adj[r.head].push_back(a.rel);
    const std::vector<std::vector<int>> comps = components(n, adj);
    std::vector<int> comp_of(n, -1);
    for (size_t c = 0; c < comps.size(); ++c)
        for (int rel : comps[c]) comp_of[rel] = (int)c;
    base.assign(n, -1);
    other.assign(n, -1);
    prim_rev.assign(n, 0);
    other_built.assign(n, 0);

    for (size_t c = 0; c < comps.size(); ++c) {
        std::vector<const Rule *> init_rules, rec_rules;
        for (const Rule &r : rules) {
            if (comp_of[r.head] != (int)c) continue;
            bool rec = false;
            for (const Atom &a : r.body) rec |= comp_of[a.rel] == (int)c;
            (rec ? rec_rules : init_rules).push_back(&r);
        }
        if (init_rules.empty() && rec_rules.empty()) continue;  // an input or empty relation

        std::vector<AOp> init_ops, rec_ops;
        int init_tmps = 0, rec_tmps = 0;
        for (const Rule *r : init_rules) {
            if (r->body.size() == 1) init_ops.push_back(plan_copy(*r, FULL));
            else plan_chain(*r, 0, FULL, &init_ops, &init_tmps);
        }
        for (const Rule *r : rec_rules)
            for (int s = 0; s < (int)r->body.size(); ++s) {
                if (comp_of[r->body[s].rel] != (int)c) continue;
                if (r->body.size() == 1) rec_ops.push_back(plan_copy(*r, DELTA));
                else plan_chain(*r, s, DELTA, &rec_ops, &rec_tmps);
            }
        // column order of each relation of the stratum: the one its
        // recursive joins read, when they read only one
        std::map<int, int> needs;
        for (const AOp &op : rec_ops) {
            if (op.kind != 0) continue;
            if (op.src_tmp < 0 && comp_of[op.src_rel] == (int)c)
                needs[op.src_rel] |= 1 << (int)op.src_rev;
            if (comp_of[op.inner_rel] == (int)c) needs[op.inner_rel] |= 1 << (int)op.inner_rev;
        }
        for (int rel : comps[c]) prim_rev[rel] = needs[rel] == 2;

        std::string names;
        for (int rel : comps[c]) names += (names.empty() ? "" : ", ") + rel_names[rel];
        auto emit = [&](const std::vector<AOp> &aops, int ntmps, bool rec, const std::string &label) {
            Step lie;
            lie.lie = true;
            lie.label = label;
            std::vector<Step> pre;
            std::vector<int> mirrors;
            auto read = [&](int rel, bool rev) {
                if (rev == (bool)prim_rev[rel]) return phys_base(rel);
                const bool mirror = rec && comp_of[rel] == (int)c;
                if (other[rel] < 0)
                    other[rel] = new_phys(rel_names[rel] + (rev ? "_2__2_1" : "_2__1_2"), rel, rev,
                                          mirror ? kMirror : kStaticRev);
                if (mirror) {
                    if (std::find(mirrors.begin(), mirrors.end(), other[rel]) == mirrors.end())
                        mirrors.push_back(other[rel]);
                } else if (!other_built[rel]) {
                    Step b;
                    b.phys = other[rel];
                    pre.push_back(b);
                }
                other_built[rel] = 1;
                return other[rel];
            };
            auto add_unique = [](std::vector<int> &v, int x) {
                if (std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
            };
            std::vector<int> tmp_phys(ntmps);
            for (int i = 0; i < ntmps; ++i) {
                tmp_phys[i] = new_phys("#tmp_" + std::to_string(steps.size()) + "_" + std::to_string(i),
                                       -1, false, kTmp);
                lie.tmps.push_back(tmp_phys[i]);
            }
            for (const AOp &a : aops) {
                Op op;
                op.kind = a.kind;
                op.src_ver = a.src_ver;
                op.pred = a.pred;
                op.map = a.map;
                if (a.kind == 1) {
                    op.src = phys_base(a.src_rel);
                    if (prim_rev[a.src_rel])
                        for (int &x : op.map.src) x = flip_source(x);
                } else {
                    op.src = a.src_tmp >= 0 ? tmp_phys[a.src_tmp] : read(a.src_rel, a.src_rev);
                    op.inner = read(a.inner_rel, a.inner_rev);
                }
                if (a.dst_tmp >= 0) {
                    op.dst = tmp_phys[a.dst_tmp];
                } else {
                    op.dst = phys_base(a.head);
                    add_unique(lie.updates, op.dst);
                    if (prim_rev[a.head]) {
                        op.map = swap_outputs(op.map);
                        if (op.pred.op != GDL_PRED_NONE) {
                            op.pred.lhs = 1 - op.pred.lhs;
                            if (!(op.pred.op & GDL_PRED_CONST)) op.pred.rhs = 1 - op.pred.rhs;
                        }
                    }
                }
                for (int ph : {op.src, op.inner})
                    if (ph >= 0 && phys[ph].kind != kTmp && comp_of[phys[ph].rel] != (int)c)
                        add_unique(lie.statics, ph);
                lie.ops.push_back(op);
            }
            if (rec) {
                for (int rel : comps[c]) add_unique(lie.updates, phys_base(rel));
                for (int mph : mirrors) {
                    Op ac;
                    ac.kind = 2;
                    ac.src = phys_base(phys[mph].rel);
                    ac.dst = mph;
                    ac.map = {2, {GDL_COL_INNER1, GDL_COL_INNER0}};
                    lie.ops.push_back(ac);
                    lie.updates.push_back(mph);
                    Step b;
                    b.phys = mph;  // the mirror starts as the reversed FULL
                    pre.push_back(b);
                }
            }
            for (Step &b : pre) {
                b.label = "build " + phys[b.phys].name;
                steps.push_back(b);
            }
            steps.push_back(lie);
        };
        if (!init_ops.empty()) emit(init_ops, init_tmps, false, names + ": init");
        if (!rec_ops.empty()) emit(rec_ops, rec_tmps, true, names + ": recursive");
    }
    for (int rel = 0; rel < n; ++rel) phys_base(rel);
}

void DatalogProgram::Impl::describe() {
    std::ostringstream os;
    auto list = [&](const char *what, const std::vector<int> &v) {
        if (v.empty()) return;
        os << " " << what << ":";
        for (int p : v) os << " " << phys[p].name;
    };
    for (const Step &s : steps) {
        if (!s.lie) {
            os << s.label << " from " << phys[phys_base(phys[s.phys].rel)].name << "\n";
            continue;
        }
        os << "lie " << s.label;
        list("update", s.updates);
        list("static", s.statics);
        list("tmp", s.tmps);
        os << "\n";
        for (const Op &op : s.ops) {
            const char *kind = op.kind == 0 ? "join" : op.kind == 1 ? "copy" : "acopy";
            os << "  " << kind << " ";
            if (op.kind == 0)
                os << phys[op.inner].name << " FULL x " << phys[op.src].name << " "
                   << ver_name(op.src_ver);
            else if (op.kind == 1)
                os << phys[op.src].name << " " << ver_name(op.src_ver);
            else
                os << phys[op.src].name << " NEWT";
            os << " -> " << phys[op.dst].name << " (" << col_name(op.map.src[0]);
            if (op.map.arity == 2) os << ", " << col_name(op.map.src[1]);
            os << ")";
            if (op.pred.op & GDL_PRED_CONST)
                os << " if c" << op.pred.lhs << " " << pred_name(op.pred.op & ~GDL_PRED_CONST) << " "
                   << (uint32_t)op.pred.rhs;
            else if (op.pred.op != GDL_PRED_NONE)
                os << " if c" << op.pred.lhs << " " << pred_name(op.pred.op) << " c" << op.pred.rhs;
            os << "\n";
        }
    }
    plan_text = os.str();
}

// ---- run --------------------------------------------------------------------

namespace {

constexpr int kGrid = 32 * 256, kBlock = 512;  // carried through the API only

// The raw rows become the sorted FULL: the fixpoint seeds delta from the
// first |FULL| raw rows, which must then be every tuple once.
void raw_from_full(Relation *r) {
    raw_rows_from_full(r);  // FULL's keys until FULL changes
    r->raw_sorted = true;
}

struct Owned {
    std::vector<Relation *> v;
    ~Owned() {
        for (Relation *r : v) {
            if (!r) continue;
            try {
                free_relation(r);
            } catch (...) {
            }
            delete r;
        }
    }
};

}  // namespace

void DatalogProgram::Impl::run(const std::map<std::string, std::vector<uint64_t>> &inputs) {
    if (!compiled) throw Error(GDL_ERR_INVALID_ARG, "datalog: the program is not compiled");
    for (const auto &kv : inputs) {
        auto it = rel_id.find(kv.first);
        if (it == rel_id.end()) throw Error(GDL_ERR_INVALID_ARG, "datalog: unknown relation " + kv.first);
        if (!is_input[it->second])
            throw Error(GDL_ERR_INVALID_ARG, "datalog: " + kv.first + " is not an .input relation");
    }
    for (size_t rel = 0; rel < rel_names.size(); ++rel) {
        if (!is_input[rel]) continue;
        auto it = inputs.find(rel_names[rel]);
        if (it == inputs.end())
            throw Error(GDL_ERR_INVALID_ARG, "datalog: no tuples given for input " + rel_names[rel]);
        if (arity[rel] == 1)
            for (uint64_t k : it->second)
                if (k & 0xffffffffull)
                    throw Error(GDL_ERR_INVALID_ARG, "datalog: unary relation " + rel_names[rel] +
                                                         ": a key with a nonzero column 1");
    }
    results.clear();
    iterations.clear();
    derived = 0;
    reset_id_bits();
    reset_checkpoint_sequence();

    Owned R;
    R.v.assign(phys.size(), nullptr);
    for (size_t p = 0; p < phys.size(); ++p) {
        const Phys &ph = phys[p];
        if (ph.kind == kMirror || ph.kind == kStaticRev) continue;  // built by their steps
        Relation *r = R.v[p] = new Relation();
        if (ph.kind == kTmp) {
            r->index_flag = false;
            load_relation(r, ph.name, 2, nullptr, 0, 1, 0, kGrid, kBlock);
            continue;
        }
        // the relation starts with its input tuples and the program's facts
        const int ar = arity[ph.rel];
        const std::vector<uint64_t> *given =
            is_input[ph.rel] ? &inputs.at(rel_names[ph.rel]) : nullptr;
        const std::vector<uint64_t> &fk = facts[ph.rel];
        const size_t n = (given ? given->size() : 0) + fk.size();
        if (n == 0) {
            load_relation(r, ph.name, ar, nullptr, 0, 1, 0, kGrid, kBlock);
            continue;
        }
        std::vector<column_type> rows(n * ar);
        for (size_t i = 0; i < n; ++i) {
            const uint64_t k = given && i < given->size() ? (*given)[i] : fk[i - (given ? given->size() : 0)];
            const column_type a = (column_type)(k >> 32);
            const column_type b = (column_type)(k & 0xffffffffull);
            if (ar == 1) {
                rows[i] = a;
            } else {
                rows[2 * i] = ph.rev ? b : a;
                rows[2 * i + 1] = ph.rev ? a : b;
            }
        }
        load_relation(r, ph.name, ar, rows.data(), n, 1, 0, kGrid, kBlock);
        raw_from_full(r);
    }
    for (const Step &s : steps) {
        if (!s.lie) {
            const Relation *src = R.v[base[phys[s.phys].rel]];
            const tuple_size_t n = src->full->tuple_counts;
            uint64_t *rev = n ? device_alloc_keys(n) : nullptr;
            Relation *r = R.v[s.phys] = new Relation();
            try {
                if (n)
                    check(gdl_project_u64(context(), src->full->keys, n,
                                          {2, {GDL_COL_INNER1, GDL_COL_INNER0}}, rev),
                          "reverse columns");
                load_relation_device(r, phys[s.phys].name, 2, rev, n, 1, 0, kGrid, kBlock);
            } catch (...) {
                device_free(rev);
                throw;
            }
            device_free(rev);
            raw_from_full(r);
            continue;
        }
        LIE lie(kGrid, kBlock);
        for (int p : s.updates) lie.add_relations(R.v[p], false);
        for (int p : s.statics) lie.add_relations(R.v[p], true);
        for (int p : s.tmps) lie.add_tmp_relation(R.v[p]);
        for (const Op &op : s.ops) {
            if (op.kind == 0)
                lie.add_ra(RelationalJoin(R.v[op.inner], FULL, R.v[op.src], op.src_ver, R.v[op.dst],
                                          op.map, op.pred));
            else if (op.kind == 1)
                lie.add_ra(RelationalCopy(R.v[op.src], op.src_ver, R.v[op.dst], op.map, op.pred));
            else
                lie.add_ra(RelationalACopy(R.v[op.src], R.v[op.dst], op.map));
        }
        lie.fixpoint_loop();
        iterations.push_back(lie.iteration_count);
        for (const IterationRecord &t : lie.trace)  // a mirror's tuples are copies
            if (phys[s.updates[t.relation]].kind != kMirror) derived += t.delta;
    }
    for (const std::string &name : kept) {
        const int rel = rel_id.at(name);
        const Relation *r = R.v[base[rel]];
        std::vector<uint64_t> keys;
        if (!prim_rev[rel]) {
            keys = relation_keys(r->full);
        } else if (other[rel] >= 0 && R.v[other[rel]]) {
            keys = relation_keys(R.v[other[rel]]->full);
        } else {
            // stored reversed: reverse on the device, sort, copy out
            const tuple_size_t n = r->full->tuple_counts;
            if (n) {
                uint64_t *rev = device_alloc_keys(n), *su = nullptr;
                tuple_size_t m = 0;
                try {
                    check(gdl_project_u64(context(), r->full->keys, n,
                                          {2, {GDL_COL_INNER1, GDL_COL_INNER0}}, rev),
                          "reverse columns");
                    m = sort_unique(rev, n, &su);
                    keys.resize(m);
                    check(gdl_memcpy_d2h(context(), keys.data(), su, m * sizeof(uint64_t)),
                          "result D2H");
                } catch (...) {
                    device_free(rev);
                    device_free(su);
                    throw;
                }
                device_free(rev);
                device_free(su);
            }
        }
        results[name] = std::move(keys);
    }
}

// ---- Slog RA plans ----------------------------------------------------------
//
// A Slog compiler plan (the shape of the reference's test/souffle/cspa.slogc)
// is one s-expression:
//   (slog-prog (<relation-decl>...) (<scc-decl>...) (<scc edge>...))
//   (relation-decl ID NAME JCOLS CANONICAL ARITY (INDEX-COLUMNS...))
//   (scc-decl NAME N RECURSIVE ((scc-rel ID DYNAMIC TMP)...) (OP...))
//   OP: (copy DST SRC VER (P...)) | (acopy DST SRC VER (P...))
//     | (join DST A VER B VER (P...))
// A relation version stores a tuple as its index columns (1-based, in order),
// then a provenance id, then the remaining columns.  A copy's or acopy's P
// index the source's stored tuple (0-based); a join's P index [the JCOLS join
// columns, A's stored columns after them, B's stored columns after them].
// The plan is read back into rules -- each join or copy is one rule over the
// relations' logical columns; acopies only maintain index versions and are
// dropped; the DELTA / FULL variants of a join give the same rule -- and
// those rules are lowered like Datalog source.  Provenance ids must not reach
// a rule head.

namespace {

struct SExp {
    std::string atom;  // empty for a list
    std::vector<SExp> list;
    int line = 0;
    bool is_list() const { return atom.empty(); }
};

SExp sexp_parse(const std::string &src) {
    constexpr int kMaxDepth = 64;  // plans nest about 6 deep
    size_t i = 0;
    int line = 1, depth = 0;
    std::function<SExp()> parse = [&]() -> SExp {
        while (i < src.size() && std::isspace((unsigned char)src[i])) line += src[i++] == '\n';
        if (i >= src.size()) fail(GDL_ERR_INVALID_ARG, line, "unexpected end of the plan");
        SExp e;
        e.line = line;
        if (src[i] == '(') {
            if (++depth > kMaxDepth) fail(GDL_ERR_INVALID_ARG, line, "the plan nests too deeply");
            ++i;
            for (;;) {
                while (i < src.size() && std::isspace((unsigned char)src[i]))
                    line += src[i++] == '\n';
                if (i >= src.size()) fail(GDL_ERR_INVALID_ARG, e.line, "unbalanced parentheses");
                if (src[i] == ')') {
                    ++i;
                    --depth;
                    return e;
                }
                e.list.push_back(parse());
            }
        }
        if (src[i] == ')') fail(GDL_ERR_INVALID_ARG, line, "unbalanced parentheses");
        const size_t j = src.find_first_of("() \t\r\n", i);
        e.atom = src.substr(i, j == std::string::npos ? std::string::npos : j - i);
        i = j == std::string::npos ? src.size() : j;
        return e;
    };
    SExp top = parse();
    while (i < src.size() && std::isspace((unsigned char)src[i])) ++i;
    if (i < src.size()) fail(GDL_ERR_INVALID_ARG, line, "text after the plan");
    return top;
}

const SExp &item(const SExp &e, size_t k, const char *what) {
    if (!e.is_list() || k >= e.list.size())
        fail(GDL_ERR_INVALID_ARG, e.line, std::string("malformed ") + what);
    return e.list[k];
}

int number(const SExp &e, const char *what) {
    if (e.is_list() || e.atom.empty() || e.atom.size() > 6 ||
        e.atom.find_first_not_of("0123456789") != std::string::npos)
        fail(GDL_ERR_INVALID_ARG, e.line, std::string("expected a small number: ") + what);
    return std::stoi(e.atom);
}

// Slog names (value-flow, $rule7-inter-body) as identifiers.
std::string slog_ident(const std::string &name) {
    std::string s;
    for (char c : name) s += std::isalnum((unsigned char)c) ? c : '_';
    if (s.empty() || std::isdigit((unsigned char)s[0])) s = "_" + s;
    return s;
}

std::string slog_to_datalog(const std::string &plan) {
    const SExp top = sexp_parse(plan);
    if (!top.is_list() || top.list.size() < 3 || item(top, 0, "plan").atom != "slog-prog")
        fail(GDL_ERR_INVALID_ARG, top.line, "not a slog-prog plan");
    struct Version {
        std::string rel;
        int jcols = 0, arity = 0;
        std::vector<int> stored;  // logical column (0-based) of each stored slot; -1: the id
    };
    std::map<std::string, Version> versions;
    std::vector<std::string> order;  // logical relations, first declared first
    std::map<std::string, int> arity;
    for (const SExp &d : item(top, 1, "relation list").list) {
        if (item(d, 0, "relation-decl").atom != "relation-decl" || d.list.size() != 7)
            fail(GDL_ERR_INVALID_ARG, d.line, "malformed relation-decl");
        Version v;
        v.rel = slog_ident(item(d, 2, "relation-decl").atom);
        v.jcols = number(item(d, 3, "relation-decl"), "join columns");
        v.arity = number(item(d, 5, "relation-decl"), "arity");
        if (v.arity < 1 || v.arity > 2)
            fail(GDL_ERR_UNSUPPORTED, d.line,
                 "relation " + v.rel + " has arity " + std::to_string(v.arity) +
                     " (unary and binary relations only)");
        std::vector<char> seen(v.arity, 0);
        for (const SExp &c : item(d, 6, "index columns").list) {
            const int col = number(c, "index column") - 1;
            if (col < 0 || col >= v.arity || seen[col])
                fail(GDL_ERR_INVALID_ARG, c.line, "bad index column");
            seen[col] = 1;
            v.stored.push_back(col);
        }
        if (v.jcols != (int)v.stored.size())
            fail(GDL_ERR_INVALID_ARG, d.line, "join columns do not match the index");
        v.stored.push_back(-1);
        for (int c = 0; c < v.arity; ++c)
            if (!seen[c]) v.stored.push_back(c);
        auto ar = arity.find(v.rel);
        if (ar == arity.end()) {
            arity[v.rel] = v.arity;
            order.push_back(v.rel);
        } else if (ar->second != v.arity) {
            fail(GDL_ERR_INVALID_ARG, d.line, "relation " + v.rel + " declared with two arities");
        }
        versions[item(d, 1, "relation-decl").atom] = v;
    }
    auto version = [&](const SExp &e) -> const Version & {
        auto it = versions.find(e.atom);
        if (it == versions.end())
            fail(GDL_ERR_INVALID_ARG, e.line, "relation version " + e.atom + " is not declared");
        return it->second;
    };
    std::set<std::string> tmp_rels, heads;
    std::vector<std::string> rules;
    std::set<std::string> rule_set;
    auto atom_text = [](const Version &v, const std::vector<std::string> &slot_vars) {
        std::vector<std::string> args(v.arity, "_");
        for (size_t k = 0; k < v.stored.size(); ++k)
            if (v.stored[k] >= 0) args[v.stored[k]] = slot_vars[k];
        std::string t = v.rel + "(";
        for (int c = 0; c < v.arity; ++c) t += (c ? ", " : "") + args[c];
        return t + ")";
    };
    for (const SExp &scc : item(top, 2, "scc list").list) {
        if (item(scc, 0, "scc-decl").atom != "scc-decl" || scc.list.size() != 6)
            fail(GDL_ERR_INVALID_ARG, scc.line, "malformed scc-decl");
        for (const SExp &r : item(scc, 4, "scc relations").list)
            if (item(r, 3, "scc-rel").atom == "#t") tmp_rels.insert(version(item(r, 1, "scc-rel")).rel);
        for (const SExp &op : item(scc, 5, "operators").list) {
            const std::string kind = item(op, 0, "operator").atom;
            if (kind == "acopy") {
                if (version(item(op, 1, "acopy")).rel != version(item(op, 2, "acopy")).rel)
                    fail(GDL_ERR_UNSUPPORTED, op.line, "an acopy between two relations");
                continue;
            }
            const Version &dst = version(item(op, 1, "operator"));
            if ((int)dst.stored.size() != dst.arity + 1 || dst.stored[dst.arity] != -1)
                fail(GDL_ERR_UNSUPPORTED, op.line, "a " + kind + " into an index version");
            std::vector<std::string> combined;  // variable of each addressable slot
            std::string body;
            int nv = 0;
            auto fresh = [&] { return "v" + std::to_string(nv++); };
            if (kind == "copy") {
                if (op.list.size() != 5) fail(GDL_ERR_INVALID_ARG, op.line, "malformed copy");
                const Version &src = version(item(op, 2, "copy"));
                for (int slot : src.stored) combined.push_back(slot < 0 ? "#id" : fresh());
                body = atom_text(src, combined);
            } else if (kind == "join") {
                if (op.list.size() != 7) fail(GDL_ERR_INVALID_ARG, op.line, "malformed join");
                const Version &a = version(item(op, 2, "join")), &b = version(item(op, 4, "join"));
                if (a.jcols != b.jcols || a.jcols < 1 || a.jcols >= a.arity + 1)
                    fail(GDL_ERR_UNSUPPORTED, op.line, "join versions with different join columns");
                std::vector<std::string> av(a.stored.size()), bv(b.stored.size());
                for (int k = 0; k < a.jcols; ++k) {
                    av[k] = bv[k] = fresh();
                    combined.push_back(av[k]);
                }
                for (size_t k = a.jcols; k < a.stored.size(); ++k)
                    combined.push_back(av[k] = a.stored[k] < 0 ? "#id" : fresh());
                for (size_t k = b.jcols; k < b.stored.size(); ++k)
                    combined.push_back(bv[k] = b.stored[k] < 0 ? "#id" : fresh());
                body = atom_text(a, av) + ", " + atom_text(b, bv);
            } else {
                fail(GDL_ERR_UNSUPPORTED, op.line, "operator " + kind + " is not supported");
            }
            const SExp &proj = item(op, op.list.size() - 1, "projection");
            if ((int)proj.list.size() != dst.arity)
                fail(GDL_ERR_INVALID_ARG, proj.line, "the projection does not match " + dst.rel);
            std::string head = dst.rel + "(";
            for (int c = 0; c < dst.arity; ++c) {
                const int p = number(item(proj, c, "projection"), "projection index");
                if (p < 0 || p >= (int)combined.size())
                    fail(GDL_ERR_INVALID_ARG, proj.line, "projection index out of range");
                if (combined[p] == "#id")
                    fail(GDL_ERR_UNSUPPORTED, proj.line, "a provenance id in a rule head");
                head += (c ? ", " : "") + combined[p];
            }
            head += ")";
            const std::string text = head + " :- " + body + ".";
            heads.insert(dst.rel);
            if (rule_set.insert(text).second) rules.push_back(text);
        }
    }
    std::ostringstream os;
    for (const std::string &rel : order) {
        os << ".decl " << rel << "(c1: number" << (arity[rel] == 2 ? ", c2: number" : "") << ")\n";
        if (!heads.count(rel)) os << ".input " << rel << "\n";
        else if (!tmp_rels.count(rel)) os << ".output " << rel << "\n";
    }
    for (const std::string &r : rules) os << r << "\n";
    return os.str();
}

}  // namespace

// ---- public -----------------------------------------------------------------

DatalogProgram::DatalogProgram() : impl_(new Impl) {}
DatalogProgram::~DatalogProgram() = default;

bool DatalogProgram::compile_slog(const std::string &plan, gdl_status *status,
                                  std::string *err) {
    std::string source;
    try {
        source = slog_to_datalog(plan);
    } catch (const CompileError &e) {
        if (status) *status = e.status;
        if (err) *err = "slog plan: " + e.msg;
        impl_.reset(new Impl);
        return false;
    }
    if (!compile(source, status, err)) {
        if (err) *err = "slog plan, rules read back: " + *err;
        return false;
    }
    return true;
}

bool DatalogProgram::compile(const std::string &source, gdl_status *status, std::string *err) {
    std::unique_ptr<Impl> fresh(new Impl);
    fresh->source = source;
    try {
        fresh->parse(lex(source));
        fresh->lower();
        fresh->describe();
    } catch (const CompileError &e) {
        if (status) *status = e.status;
        if (err) *err = e.msg;
        impl_.reset(new Impl);
        return false;
    }
    for (size_t rel = 0; rel < fresh->rel_names.size(); ++rel) {
        if (fresh->any_kept ? fresh->is_kept[rel] : !fresh->is_input[rel] && !fresh->hidden[rel])
            fresh->kept.push_back(fresh->rel_names[rel]);
        if (fresh->is_input[rel]) fresh->input_names.push_back(fresh->rel_names[rel]);
    }
    fresh->compiled = true;
    impl_ = std::move(fresh);
    if (status) *status = GDL_OK;
    if (err) err->clear();
    return true;
}

void DatalogProgram::run(const std::map<std::string, std::vector<uint64_t>> &inputs) {
    impl_->run(inputs);
}

const std::vector<uint64_t> *DatalogProgram::result(const std::string &relation) const {
    auto it = impl_->results.find(relation);
    return it == impl_->results.end() ? nullptr : &it->second;
}

const std::vector<std::string> &DatalogProgram::kept() const { return impl_->kept; }
const std::vector<std::string> &DatalogProgram::inputs() const { return impl_->input_names; }

const std::string &DatalogProgram::input_file(const std::string &relation) const {
    static const std::string none;
    auto it = impl_->rel_id.find(relation);
    if (it == impl_->rel_id.end()) return none;
    auto f = impl_->input_file.find(it->second);
    return f == impl_->input_file.end() ? none : f->second;
}

int DatalogProgram::arity(const std::string &relation) const {
    auto it = impl_->rel_id.find(relation);
    return it == impl_->rel_id.end() ? 0 : impl_->arity[it->second];
}

const std::string &DatalogProgram::plan() const { return impl_->plan_text; }
const std::string &DatalogProgram::source() const { return impl_->source; }
const std::vector<int> &DatalogProgram::iterations() const { return impl_->iterations; }
uint64_t DatalogProgram::derived() const { return impl_->derived; }

}  // namespace gdlog

// ---- C ABI ------------------------------------------------------------------

struct gdl_datalog {
    gdlog::DatalogProgram prog;
};

namespace {
thread_local std::string g_dl_error;
}

extern "C" {

gdl_status gdl_datalog_compile(const char *source, gdl_datalog **out) {
    g_dl_error.clear();
    if (!source || !out) return GDL_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<gdl_datalog> p(new gdl_datalog);
    gdl_status st = GDL_OK;
    try {
        if (!p->prog.compile(source, &st, &g_dl_error)) return st;
    } catch (const std::exception &e) {
        g_dl_error = e.what();
        return GDL_ERR_INVALID_ARG;
    }
    *out = p.release();
    return GDL_OK;
}

gdl_status gdl_datalog_compile_slog(const char *plan, gdl_datalog **out) {
    g_dl_error.clear();
    if (!plan || !out) return GDL_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<gdl_datalog> p(new gdl_datalog);
    gdl_status st = GDL_OK;
    try {
        if (!p->prog.compile_slog(plan, &st, &g_dl_error)) return st;
    } catch (const std::exception &e) {
        g_dl_error = e.what();
        return GDL_ERR_INVALID_ARG;
    }
    *out = p.release();
    return GDL_OK;
}

void gdl_datalog_destroy(gdl_datalog *prog) { delete prog; }

const char *gdl_datalog_source(const gdl_datalog *prog) {
    return prog ? prog->prog.source().c_str() : "";
}

const char *gdl_datalog_plan(const gdl_datalog *prog) {
    return prog ? prog->prog.plan().c_str() : "";
}

gdl_status gdl_datalog_run(gdl_ctx *ctx, gdl_datalog *prog, uint32_t n_inputs,
                           const char *const *names, const uint64_t *const *keys,
                           const uint64_t *counts) {
    g_dl_error.clear();
    if (!ctx || !prog || (n_inputs && (!names || !keys || !counts))) return GDL_ERR_INVALID_ARG;
    std::map<std::string, std::vector<uint64_t>> in;
    for (uint32_t i = 0; i < n_inputs; ++i) {
        if (!names[i] || (counts[i] && !keys[i])) return GDL_ERR_INVALID_ARG;
        std::vector<uint64_t> &v = in[names[i]];
        v.insert(v.end(), keys[i], keys[i] + counts[i]);
    }
    gdl_ctx *prev = gdlog::context_or_null();
    gdlog::set_context(ctx);
    gdl_status st = GDL_OK;
    try {
        prog->prog.run(in);
    } catch (const gdlog::Error &e) {
        g_dl_error = e.what();
        st = e.status;
    } catch (const std::exception &e) {
        g_dl_error = e.what();
        st = GDL_ERR_INVALID_ARG;
    }
    gdlog::set_context(prev);
    return st;
}

gdl_status gdl_datalog_result(const gdl_datalog *prog, const char *relation, uint64_t *keys,
                              uint64_t cap, uint64_t *n) {
    if (!prog || !relation || !n) return GDL_ERR_INVALID_ARG;
    const std::vector<uint64_t> *r = prog->prog.result(relation);
    if (!r) {
        g_dl_error = std::string("datalog: no result for ") + relation;
        return GDL_ERR_INVALID_ARG;
    }
    *n = r->size();
    if (!keys) return GDL_OK;
    if (r->size() > cap) return GDL_ERR_OVERFLOW;
    if (!r->empty()) std::memcpy(keys, r->data(), r->size() * sizeof(uint64_t));
    return GDL_OK;
}

gdl_status gdl_datalog_stats(const gdl_datalog *prog, uint64_t *derived, int32_t *iterations,
                             uint32_t cap, uint32_t *n_lies) {
    if (!prog) return GDL_ERR_INVALID_ARG;
    const std::vector<int> &it = prog->prog.iterations();
    if (derived) *derived = prog->prog.derived();
    if (n_lies) *n_lies = (uint32_t)it.size();
    if (!iterations) return GDL_OK;
    if (it.size() > cap) return GDL_ERR_OVERFLOW;
    for (size_t i = 0; i < it.size(); ++i) iterations[i] = it[i];
    return GDL_OK;
}

const char *gdl_datalog_kept(const gdl_datalog *prog, uint32_t i) {
    if (!prog) return nullptr;
    const std::vector<std::string> &k = prog->prog.kept();
    return i < k.size() ? k[i].c_str() : nullptr;
}

int gdl_datalog_arity(const gdl_datalog *prog, const char *relation) {
    return prog && relation ? prog->prog.arity(relation) : 0;
}

const char *gdl_datalog_last_error(void) { return g_dl_error.c_str(); }

}  // extern "C"
