// In-launch cones of the four-level cut schedule (k_cuts_quad).  Kept apart
// from the cut kernels so that a host build can check them serially.
#pragma once

// The cone of node v in a launch whose levels hold ids [s[0], s[1]),
// [s[1], s[2]), [s[2], s[3]), [s[3], e): the nodes of each of the three lower
// levels that v reaches inside the launch (at most 8, 4 and 2: a cone is a
// binary tree at most four levels high), each with its fanin literals.
struct Cone4 {
    uint32_t id[3][8];     // id[p][k]: k-th cone node on level p of the launch
    uint32_t l0[3][8], l1[3][8];
    uint32_t n[3];
};

__device__ __forceinline__ int cone4_level(uint32_t x, const uint32_t *s) {
    return x < s[0] ? -1 : x < s[1] ? 0 : x < s[2] ? 1 : x < s[3] ? 2 : 3;
}

// Adds x to its level's list unless it is below the launch or listed already.
__device__ __forceinline__ void cone4_add(Cone4 &c, uint32_t x, const uint32_t *s, const uint32_t *__restrict__ f0,
                                          const uint32_t *__restrict__ f1) {
    const int p = cone4_level(x, s);
    if (p < 0 || p > 2) return;
    for (uint32_t k = 0; k < c.n[p]; k++)
        if (c.id[p][k] == x) return;
    const uint32_t k = c.n[p];
    if (k >= 8u) return;    // cannot happen: the lists are bounded by the tree
    c.id[p][k] = x;
    c.l0[p][k] = f0[x];
    c.l1[p][k] = f1[x];
    c.n[p] = k + 1;
}

// Fills c for node v on level pv (0..3) of the launch: its fanins, then the
// fanins of the listed nodes, top level first.
__device__ __forceinline__ void cone4_discover(uint32_t v, const uint32_t *s, const uint32_t *__restrict__ f0,
                                               const uint32_t *__restrict__ f1, Cone4 &c) {
    c.n[0] = c.n[1] = c.n[2] = 0;
    cone4_add(c, LID(f0[v]), s, f0, f1);
    cone4_add(c, LID(f1[v]), s, f0, f1);
    for (int p = 2; p >= 1; p--)
        for (uint32_t k = 0; k < c.n[p]; k++) {
            cone4_add(c, LID(c.l0[p][k]), s, f0, f1);
            cone4_add(c, LID(c.l1[p][k]), s, f0, f1);
        }
}
