// gpuls: the standalone shell of the rw/strash path over libgls.so.
//
// Mirrors the reference's front end (src/main.cpp:56-83) and command registry
// (src/command_manager.cpp:254-306): `gpuls -c "read in.aig; st; rw; rwz; ps;
// write out.aig"` runs a script, `gpuls` alone reads commands from stdin.
// Handlers have the reference's signature and return convention: 0 = ok,
// 1 = not recognised / wrong format.  `rwz` is registered as `rw -z`
// (the reference has no rwz command).
#include "gls.h"
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Shell {
    gls_ctx *ctx = nullptr;
    gls_aig *aig = nullptr;
    int verbose = 1;
    double prevAlg = 0, prevFull = 0, totalAlg = 0, totalFull = 0;
};

using CommandHandler = std::function<int(Shell &, const std::vector<std::string> &)>;

bool has_flag(const std::vector<std::string> &v, const char *f) {
    for (size_t i = 1; i < v.size(); i++)
        if (v[i] == f) return true;
    return false;
}

// flags of the form -z -d -s -c -v N; anything else is a format error
bool flags_ok(const std::vector<std::string> &v, const char *allowed) {
    for (size_t i = 1; i < v.size(); i++) {
        const std::string &a = v[i];
        if (a == "-v") { if (++i >= v.size()) return false; continue; }
        if (a.size() != 2 || a[0] != '-' || !strchr(allowed, a[1])) return false;
    }
    return true;
}

int need_aig(Shell &sh, const char *cmd) {
    if (!sh.aig) { printf("%s: AIG is null! \n", cmd); return 0; }
    return 1;
}

void report_error(Shell &sh, const char *what, int rc) {
    printf("%s: failed (%d): %s\n", what, rc, gls_last_error(sh.ctx));
}

template <class F> int timed(Shell &sh, F f) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = f();
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    sh.prevAlg = sh.prevFull = dt;
    sh.totalAlg += dt;
    sh.totalFull += dt;
    return rc;
}

int readHandler(Shell &sh, const std::vector<std::string> &v) {
    if (v.size() < 2) return 1;
    gls_aig *a = nullptr;
    int rc = gls_aig_read_aiger(sh.ctx, v[1].c_str(), &a);
    if (rc != GLS_OK) { report_error(sh, "read", rc); return 0; }
    if (sh.aig) gls_aig_free(sh.aig);
    sh.aig = a;
    return 0;
}

int writeHandler(Shell &sh, const std::vector<std::string> &v) {
    if (v.size() < 2) return 1;
    const std::string &p = v[1];
    if (p.size() < 4 || p.compare(p.size() - 4, 4, ".aig") != 0) { printf("write: only support .aig format!\n"); return 1; }
    if (!need_aig(sh, "saveFile")) return 0;
    int rc = gls_aig_write_aiger(sh.aig, p.c_str());
    if (rc != GLS_OK) report_error(sh, "write", rc);
    else printf("Output AIG file saved at path: %s\n", p.c_str());
    return 0;
}

int psHandler(Shell &sh, const std::vector<std::string> &) {
    if (!need_aig(sh, "ps")) return 0;
    gls_stats s;
    gls_aig_stats(sh.aig, &s);
    printf("AIG stats: i/o = %d/%d and = %d level = %d\n", s.nPIs, s.nPOs, s.nAnds, s.levels);
    return 0;
}

int timeHandler(Shell &sh, const std::vector<std::string> &) {
    printf("{time} prev cmd: alg %.2lf s, full %.2lf s; total: alg %.2lf s, full %.2lf s.\n", sh.prevAlg, sh.prevFull,
           sh.totalAlg, sh.totalFull);
    return 0;
}

int strashHandler(Shell &sh, const std::vector<std::string> &v) {
    if (!flags_ok(v, "cv")) return 1;
    if (!need_aig(sh, "strash")) return 0;
    const int32_t flags = has_flag(v, "-c") ? GLS_ST_CPU : 0;
    int rc = timed(sh, [&] { return gls_strash(sh.aig, flags); });
    if (rc != GLS_OK) report_error(sh, "strash", rc);
    return 0;
}

int rewriteHandler(Shell &sh, const std::vector<std::string> &v) {
    if (!flags_ok(v, "zdv")) return 1;
    if (!need_aig(sh, "rewrite")) return 0;
    bool z = has_flag(v, "-z") || v[0] == "rwz";
    int rc = timed(sh, [&] { return gls_rewrite(sh.aig, z ? 1 : 0, 0); });
    if (rc != GLS_OK) { report_error(sh, "rewrite", rc); return 0; }
    if (sh.verbose >= 1) {
        gls_op_info info;
        gls_aig_last_op(sh.aig, &info);
        gls_stats s;
        gls_aig_stats(sh.aig, &s);
        printf("rewrite: %s%lld candidates, %lld replaced, %lld new nodes; and = %d, %.2f ms\n", z ? "(zeros) " : "",
               (long long)info.candidates, (long long)info.selected, (long long)info.new_nodes, s.nAnds, info.total_ms);
    }
    return 0;
}

int balanceHandler(Shell &sh, const std::vector<std::string> &v) {
    if (!flags_ok(v, "sv")) return 1;
    if (!need_aig(sh, "balance")) return 0;
    // -s: ties of equal level in thrust's shuffle order instead of by id
    // (command_manager.cpp:61-74)
    const int sortDecId = has_flag(v, "-s") ? 0 : 1;
    if (!sortDecId) printf("** balance without using id as tie break when sorting.\n");
    // balance, then strash to update levels (command_manager.cpp:77)
    int rc = timed(sh, [&] { return gls_balance(sh.aig, sortDecId, 1); });
    if (rc != GLS_OK) report_error(sh, "balance", rc);
    return 0;
}

struct CmdMan {
    std::map<std::string, CommandHandler> table;
    void registerCommand(const std::string &name, CommandHandler h) { table[name] = std::move(h); }
    void registerAll() {
        registerCommand("read", readHandler);
        registerCommand("write", writeHandler);
        registerCommand("ps", psHandler);
        registerCommand("print_stats", psHandler);
        registerCommand("time", timeHandler);
        registerCommand("st", strashHandler);
        registerCommand("strash", strashHandler);
        registerCommand("rw", rewriteHandler);
        registerCommand("rewrite", rewriteHandler);
        registerCommand("rwz", rewriteHandler);
        registerCommand("b", balanceHandler);
        registerCommand("balance", balanceHandler);
    }
    int launch(Shell &sh, const std::vector<std::string> &v) {
        auto it = table.find(v[0]);
        if (it == table.end()) {
            printf("Command %s not registered, ignored.\n", v[0].c_str());
            return 1;
        }
        int ret = it->second(sh, v);
        if (ret == 1) printf("Command %s not recognized or in wrong format!\n", v[0].c_str());
        return ret;
    }
};

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

int run_line(CmdMan &cm, Shell &sh, const std::string &cmd) {
    std::string c = trim(cmd);
    if (c.empty()) return 0;
    std::istringstream in(c);
    std::vector<std::string> v;
    std::string w;
    while (in >> w) v.push_back(w);
    return cm.launch(sh, v);
}

} // namespace

int main(int argc, char **argv) {
    std::string script;
    int device = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if ((a == "-c" || a == "--script") && i + 1 < argc) script = argv[++i];
        else if (a == "--device" && i + 1 < argc) device = atoi(argv[++i]);
        else { fprintf(stderr, "usage: gpuls [-c \"cmd; cmd\"] [--device N]\n"); return 1; }
    }
    Shell sh;
    int rc = gls_ctx_create(device, &sh.ctx);
    if (rc != GLS_OK) {
        fprintf(stderr, "gpuls: no usable GPU (%d); only read/write/ps work\n", rc);
        gls_ctx_create(-1, &sh.ctx);
    }
    CmdMan cm;
    cm.registerAll();
    if (!script.empty()) {
        printf("Running in script mode: \"%s\".\n============================================\n", script.c_str());
        size_t start = 0;
        while (start <= script.size()) {
            size_t end = script.find(';', start);
            if (end == std::string::npos) end = script.size();
            run_line(cm, sh, script.substr(start, end - start));
            start = end + 1;
        }
    } else {
        std::string line;
        printf("gpuls> ");
        fflush(stdout);
        while (std::getline(std::cin, line)) {
            std::string t = trim(line);
            if (t == "exit" || t == "quit") break;
            run_line(cm, sh, t);
            printf("gpuls> ");
            fflush(stdout);
        }
    }
    if (sh.aig) gls_aig_free(sh.aig);
    gls_ctx_destroy(sh.ctx);
    return 0;
}
