// khs_layout.h — state array ids shared by host and device code (no HIP headers).
#pragma once

namespace khs {

// Evolved state: 12 field + 2x6 fluid arrays, reference layout x + z*nx.
// The order matches enum khs_array of include/khs.h.
enum StateId {
    S_EX1 = 0, S_EX2, S_EX3, S_EZ1, S_EZ2, S_EZ3, S_BY1, S_BY2, S_BY3, S_AZ1, S_AZ2, S_AZ3,
    S_NA, S_PXA, S_PZA, S_Q1A, S_Q2A, S_Q3A,
    S_NB, S_PXB, S_PZB, S_Q1B, S_Q2B, S_Q3B,
    S_COUNT
};
// fluid variable order inside one species block
enum FluidVar { F_N = 0, F_PX, F_PZ, F_Q1, F_Q2, F_Q3, F_COUNT };

}  // namespace khs
