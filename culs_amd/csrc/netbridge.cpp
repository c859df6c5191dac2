// gls_aig_from_net: an in-process AIG network (objects with fanin indices,
// as ABC's strashed networks hold them) to a gls_aig, numbered as the
// reference's ABC patch numbers it when it hands a network to the GPU side
// (AbcNtkToGpuMan, src/abc_patch/abc_patch_transform.cpp:5-87): PIs in the
// network's CI order, ANDs in depth-first post-order from the COs (Abc_NtkDfs:
// COs in order, fanin 0 before fanin 1, each node once, unreachable nodes
// left out), the constant-1 node as the constant.  Host code; the AIG is
// uploaded on first use like any other.
#include "gls_internal.h"

#include <vector>

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

} // namespace

extern "C" int gls_aig_from_net(gls_ctx *ctx, const gls_net *net, gls_aig **out) {
    if (!ctx || !net || !out) return GLS_ERR_ARG;
    *out = nullptr;
    const int32_t n = net->nObjs;
    if (n < 0 || net->nCis < 0 || net->nCos < 0 || (n > 0 && (!net->kind || !net->fanin0 || !net->fanin1 ||
                                                              !net->compl0 || !net->compl1)) ||
        (net->nCis > 0 && !net->ci) || (net->nCos > 0 && !net->co))
        return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: bad arguments");
    for (int32_t v = 0; v < n; v++)
        if (net->kind[v] > GLS_NET_CO) return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: object %d has kind %d", v, net->kind[v]);
    // literal of each object once numbered (AIGER convention); CIs first
    std::vector<uint32_t> lit(n, NONE);
    for (int32_t k = 0; k < net->nCis; k++) {
        const int32_t v = net->ci[k];
        if (v < 0 || v >= n || net->kind[v] != GLS_NET_CI || lit[v] != NONE)
            return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: CI entry %d (object %d) is not a distinct CI", k, v);
        lit[v] = 2u * (uint32_t)(k + 1);
    }
    for (int32_t v = 0; v < n; v++)
        if (net->kind[v] == GLS_NET_CONST1) lit[v] = 1u;   // constant TRUE
    auto fanin_ok = [&](int32_t f) {
        return f >= 0 && f < n && (net->kind[f] == GLS_NET_AND || net->kind[f] == GLS_NET_CONST1 || lit[f] != NONE);
    };
    // every AND's and CO's fanins, reached or not
    for (int32_t v = 0; v < n; v++) {
        if (net->kind[v] == GLS_NET_AND && (!fanin_ok(net->fanin0[v]) || !fanin_ok(net->fanin1[v])))
            return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: AND %d has an invalid fanin", v);
        if (net->kind[v] == GLS_NET_CO && !fanin_ok(net->fanin0[v]))
            return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: CO object %d has an invalid fanin", v);
    }
    // depth-first post-order from the COs; state 1 = on the stack, 2 = done
    const int32_t nPIs = net->nCis;
    std::vector<uint8_t> state(n, 0);
    std::vector<int32_t> order;
    std::vector<std::pair<int32_t, int32_t>> stack;   // (object, next fanin to visit)
    for (int32_t c = 0; c < net->nCos; c++) {
        const int32_t co = net->co[c];
        if (co < 0 || co >= n || net->kind[co] != GLS_NET_CO)
            return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: CO entry %d (object %d) is not a CO", c, co);
        const int32_t root = net->fanin0[co];
        if (net->kind[root] != GLS_NET_AND || state[root] == 2) continue;
        stack.push_back({root, 0});
        state[root] = 1;
        while (!stack.empty()) {
            auto &top = stack.back();
            const int32_t v = top.first;
            if (top.second < 2) {
                const int32_t f = top.second == 0 ? net->fanin0[v] : net->fanin1[v];
                ++top.second;
                if (net->kind[f] != GLS_NET_AND || state[f] == 2) continue;
                if (state[f] == 1) return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_net: a cycle through object %d", f);
                state[f] = 1;
                stack.push_back({f, 0});
                continue;
            }
            state[v] = 2;
            lit[v] = 2u * (uint32_t)(nPIs + 1 + (int32_t)order.size());
            order.push_back(v);
            stack.pop_back();
        }
    }
    if ((int64_t)nPIs + 1 + (int64_t)order.size() > 0x7FFFFFF0LL) return gls_fail(ctx, GLS_ERR_ARG, "AIG too large");
    const int32_t nAnds = (int32_t)order.size(), nObjs = nPIs + 1 + nAnds;
    std::vector<uint32_t> f0(nObjs, 0), f1(nObjs, 0), po(net->nCos);
    for (int32_t k = 0; k < nAnds; k++) {
        const int32_t v = order[k];
        f0[nPIs + 1 + k] = lit[net->fanin0[v]] ^ (net->compl0[v] ? 1u : 0u);
        f1[nPIs + 1 + k] = lit[net->fanin1[v]] ^ (net->compl1[v] ? 1u : 0u);
    }
    for (int32_t c = 0; c < net->nCos; c++) {
        const int32_t co = net->co[c];
        po[c] = lit[net->fanin0[co]] ^ (net->compl0[co] ? 1u : 0u);
    }
    return gls_aig_from_arrays(ctx, nPIs, net->nCos, nAnds, f0.data(), f1.data(), po.empty() ? nullptr : po.data(),
                               GLS_LIT_AIGER, out);
}
