// Variable placement in 4-input truth tables (16 bits) by delta swaps.
//
// The merge of two cuts places a fanin cut's variables at the positions its
// leaves take in the merged leaf list (the set bits of `phase`), and the
// support reduction moves the variables a table depends on (the set bits of
// `phase`) down to positions 0..k-1.  The reference does both with chains of
// adjacent variable swaps (rewrite.cu:203-266).  A cut's table over k leaves
// does not depend on variables k..3, so both are permutations of the index
// bits that only move a variable past variables the table does not depend
// on, and the resulting 16-bit table is the same whatever permutation
// realises the placement:
//   - placement (stretch): for j = k-1 down to 0, swap variable j with
//     variable p_j, the position of the (j+1)-th set bit of phase.  p_j >= j,
//     and when p_j > j the variable at p_j is one the table does not depend
//     on at that point (k..3 at first, then the positions vacated above);
//   - reduction (shrink): the same swaps in the opposite order.
// Swapping index bits j < p is one delta swap: the entries whose index has
// bit j set and bit p clear exchange with the entries d = 2^p - 2^j above.
// At most three swaps are needed (j <= 2; p_3 = 3 whenever k = 4).
//
// Entry of a phase: up to three swaps in placement order, 20 bits each (the
// 16-bit mask of the lower entries, then d in 4 bits); an unused swap has
// mask 0 and changes nothing.
#pragma once

#include <cstdint>

constexpr uint32_t tt_swap_mask(int j, int p) {
    uint32_t m = 0;
    for (int q = 0; q < 16; q++)
        if (((q >> j) & 1) && !((q >> p) & 1)) m |= 1u << q;
    return m;
}

constexpr uint64_t tt_place_entry(int phase) {
    int pos[4] = {0, 0, 0, 0};
    int k = 0;
    for (int i = 0; i < 4; i++)
        if ((phase >> i) & 1) pos[k++] = i;
    uint64_t e = 0;
    int s = 0;
    for (int j = k - 1; j >= 0; j--) {
        if (pos[j] == j) break;     // then every lower variable stays as well
        const uint64_t d = (1u << pos[j]) - (1u << j);
        e |= ((uint64_t)tt_swap_mask(j, pos[j]) | d << 16) << (20 * s);
        s++;
    }
    return e;
}

__host__ __device__ __forceinline__ uint32_t tt_delta_swap(uint32_t t, uint64_t e, int s) {
    const uint32_t m = (uint32_t)(e >> (20 * s)) & 0xFFFFu;
    const uint32_t d = (uint32_t)(e >> (20 * s + 16)) & 15u;
    const uint32_t x = ((t >> d) ^ t) & m;
    return t ^ x ^ (x << d);
}

// the table over k leaves placed at the set bits of the entry's phase
__host__ __device__ __forceinline__ uint32_t tt_place_stretch(uint32_t t, uint64_t e) {
    t = tt_delta_swap(t, e, 0);
    t = tt_delta_swap(t, e, 1);
    return tt_delta_swap(t, e, 2);
}

// the variables at the set bits of the entry's phase moved down to 0..k-1
__host__ __device__ __forceinline__ uint32_t tt_place_shrink(uint32_t t, uint64_t e) {
    t = tt_delta_swap(t, e, 2);
    t = tt_delta_swap(t, e, 1);
    return tt_delta_swap(t, e, 0);
}
