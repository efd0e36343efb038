// frontend_asan_main.cpp -- compiles Datalog programs (no device run) through
// the front end's C ABI, for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of the front end's host code (`make sanitize`).  Test infrastructure
// only: the programs and malformed inputs exercise the parser, the
// stratifier and the lowering, including their error paths.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "gdl_datalog.h"

static std::string slurp(const char *path) {
    std::ifstream f(path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

int main(int argc, char **argv) {
    int bad = 0, n = 0;
    for (int i = 1; i < argc; ++i) {  // program files that must compile
        gdl_datalog *p = nullptr;
        const std::string src = slurp(argv[i]);
        const bool slog = src.find("(") == 0 || std::string(argv[i]).find(".slogc") != std::string::npos;
        gdl_status st = slog ? gdl_datalog_compile_slog(src.c_str(), &p)
                             : gdl_datalog_compile(src.c_str(), &p);
        if (st != GDL_OK) {
            std::fprintf(stderr, "%s: %s\n", argv[i], gdl_datalog_last_error());
            ++bad;
        } else {
            (void)gdl_datalog_plan(p);
            gdl_datalog_destroy(p);
        }
        ++n;
    }
    const char *broken[] = {
        "", ".decl", ".decl e(a:number", "e(1,2).", ".decl e(a:number,b:number)\ne(x,y) :- e(x,",
        ".decl a(x:number)\n.decl b(x:number)\na(x) :- !b(x).\nb(x) :- !a(x).",
        ".decl p(x:number,y:number)\np(x,y) :- p(y,x), x != 99999999999999999999999.",
        ".decl p(x:number)\np(x) :- ((((((((((((((((((((((((((((((((p(x)))))))))))))))))))))))))))))))).",
        ".decl e(a:number,b:number)\n.decl t(a:number,b:number)\nt(x,z) :- e(x,y), e(y,w), e(w,v), e(v,u), e(u,z).",
    };
    for (const char *s : broken) {  // must fail cleanly or compile, never crash
        gdl_datalog *p = nullptr;
        if (gdl_datalog_compile(s, &p) == GDL_OK) gdl_datalog_destroy(p);
        ++n;
    }
    std::printf("front-end sanitizer run: %d programs, %s\n", n, bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
