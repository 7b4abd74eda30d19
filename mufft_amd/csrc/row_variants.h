// Registered variants of the fused row kernel (host and device).
//
// Each entry (log2 N, variant) is instantiated for c2c forward/inverse, r2c
// and c2r.  The planner's static choice is rv_default(N) (geometry.h); the
// autotuner times the other entries of a size on the device and may pick one.
#pragma once

#include "geometry.h"

namespace mufft {

struct RowVariantEntry {
    int logN;
    int var;
};

// Shorthands: E points per thread, exchange (P padded, S split, F flat),
// twiddles (full table or split product).
constexpr int V16P = rv_make(16, 0, 0);
constexpr int V16PT = rv_make(16, 0, 1);
constexpr int V16S = rv_make(16, 1, 0);
constexpr int V16F = rv_make(16, 2, 0);
constexpr int V8P = rv_make(8, 0, 0);
constexpr int V32P = rv_make(32, 0, 0);
constexpr int V32PT = rv_make(32, 0, 1);
constexpr int V32S = rv_make(32, 1, 0);
constexpr int V32ST = rv_make(32, 1, 1);
constexpr int V64ST = rv_make(64, 1, 1);
constexpr int V64PT = rv_make(64, 0, 1);
// Register budgets for two workgroups per CU (LDS must allow it too).
constexpr int V16ST = rv_make(16, 1, 1);
constexpr int V16ST_2 = rv_with_occ(V16ST, 2);
constexpr int V16S_2 = rv_with_occ(V16S, 2);
constexpr int V32ST_2 = rv_with_occ(V32ST, 2);
constexpr int V16PT_2 = rv_with_occ(V16PT, 2);
constexpr int V16S_4 = rv_with_occ(V16S, 4);
// 32 points per thread in radix-16 passes (lower register peak per butterfly).
constexpr int V32ST_R16 = rv_with_radix(V32ST, 4);
constexpr int V32PT_R16 = rv_with_radix(V32PT, 4);
// c2r prologue loads in ordered chunks (the other kinds ignore the bit).
constexpr int V32ST_C = rv_with_chunked_loads(V32ST);
constexpr int V32PT_C = rv_with_chunked_loads(V32PT);
constexpr int V16PT_C = rv_with_chunked_loads(V16PT);
// Two workgroups per transform (split_impl.h), with the sub-transform's variant.
constexpr int S2_16PT = rv_make_split(1, V16PT);
constexpr int S2_32PT = rv_make_split(1, V32PT);
constexpr int S2_32ST = rv_make_split(1, V32ST);
constexpr int S2_32ST_2 = rv_make_split(1, V32ST_2);
constexpr int S2_16S_2 = rv_make_split(1, V16S_2);
constexpr int S2_16ST_2 = rv_make_split(1, V16ST_2);
constexpr int S2_32ST_C = rv_with_chunked_loads(S2_32ST);
constexpr int S2_32ST_2C = rv_with_chunked_loads(S2_32ST_2);
constexpr int V32ST_2C = rv_with_chunked_loads(V32ST_2);
// c2r bins read once through mirrored lanes (the other kinds ignore the bit).
constexpr int V16PT_M = rv_with_mirror_lanes(V16PT);
constexpr int V32PT_M = rv_with_mirror_lanes(V32PT);
constexpr int V32ST_M = rv_with_mirror_lanes(V32ST);
constexpr int V32ST_2M = rv_with_mirror_lanes(V32ST_2);
constexpr int S2_32ST_M = rv_with_mirror_lanes(S2_32ST);
constexpr int S2_32ST_2M = rv_with_mirror_lanes(S2_32ST_2);

constexpr RowVariantEntry kRowVariants[] = {
    {1, V16P}, {2, V16P}, {3, V16P}, {4, V16P},
    {5, V16P}, {5, V8P},
    {6, V16P}, {6, V8P},
    {7, V16P}, {7, V8P},
    {8, V16P}, {8, V8P},
    {9, V16P}, {9, V16PT}, {9, V8P}, {9, V32PT},
    {10, V16P}, {10, V16PT}, {10, V8P}, {10, V16S}, {10, V16F}, {10, V32PT}, {10, V16PT_M},
    {11, V16P}, {11, V16PT}, {11, V8P}, {11, V16S}, {11, V16F}, {11, V32PT}, {11, V16PT_M},
    {12, V16P}, {12, V16PT}, {12, V8P}, {12, V16S}, {12, V16F}, {12, V32PT}, {12, V16PT_M},
    {13, V32P}, {13, V32PT}, {13, V16PT}, {13, V16S}, {13, V16S_4}, {13, V16PT_2}, {13, V16PT_C},
    {13, V16PT_M},
    {14, V32P}, {14, V32PT}, {14, V16PT}, {14, V64PT}, {14, V32ST_2}, {14, V16S_2}, {14, V16ST_2},
    {14, V32PT_R16}, {14, V32PT_C}, {14, V32ST_2C}, {14, V32PT_M}, {14, V32ST_2M},
    {14, S2_16PT},
    {15, V32S}, {15, V32ST}, {15, V64ST}, {15, S2_32PT}, {15, S2_16PT}, {15, S2_32ST},
    {15, S2_32ST_2}, {15, S2_16S_2}, {15, S2_16ST_2}, {15, V32ST_R16}, {15, V32ST_C},
    {15, S2_32ST_C}, {15, S2_32ST_2C}, {15, V32ST_M}, {15, S2_32ST_M}, {15, S2_32ST_2M},
    {16, S2_32ST},
};
constexpr int kNumRowVariants = sizeof(kRowVariants) / sizeof(kRowVariants[0]);

// The instantiation units split the registry at the first entry of 8192 and
// of 32768 points (the entries are ordered by size; the larger sizes take
// most of the compile time).
constexpr int row_split_index(int logn, int i = 0) {
    return i >= kNumRowVariants || kRowVariants[i].logN >= logn ? i : row_split_index(logn, i + 1);
}
constexpr int kRowMidIndex = row_split_index(13);
constexpr int kRowSplitIndex = row_split_index(15);

}  // namespace mufft
