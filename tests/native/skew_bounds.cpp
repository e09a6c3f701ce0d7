// Enumerates every index the lexicographic sweep (gsLexWave, simple_fp64.hip)
// and the writers of its skewed layout (k_mom, k_sst_asm, k_gs_unskew) form,
// for a set of mesh shapes, block widths and read-ahead depths, and checks
// each against its allocation: the device arrays of skewLength, the LDS
// arrays (row 0, the hand-off ring, the dummy slots, the partner map, the
// step counters). It also checks that every fetch of a lane on the wavefront
// reads the cell the step needs, that a store off the wavefront lands in a
// slot of no cell or a parking slot, and that no fetch reads a parking slot.
// Mirrors the index expressions of gsLexWave. Also checks the colour-packed
// layout of the mixed path's two-colour sweeps. Prints "ok" or the first
// violation.
#include <algorithm>
#include <cstdio>
#include <set>

#include "../../cfd-simulator_amd/csrc/device/layout.hpp"

using namespace cfdgpu;

static int fail(const char* what, int nI, int nJ, int B, long v, long lim) {
  std::printf("FAIL %s nI=%d nJ=%d B=%d index=%ld limit=%ld\n", what, nI, nJ, B, v, lim);
  return 1;
}

// B: the block's lane count (>= R when the instance is one of a batch whose
// widest member has more rows); lanes past R fetch as lane R - 1 and park
// their stores.
static int check(int nI, int nJ, int D, int B) {
  const int R = gsLanes(nJ);
  if (B < R) return 0;
  if (R < nJ) return fail("lanes < rows", nI, nJ, B, R, nJ);
  if (D > kSkewFront || 2 * D + 2 > kSkewTail || B - R > kSkewPark) return fail("padding", nI, nJ, B, D, 0);
  const long len = skewLength(nI, nJ, R);
  const int steps = skewSteps(nI, nJ);
  const long parkStart = static_cast<long>(kSkewFront + steps + kSkewTail) * R;
  const int W = B / 64;
  // writers: one slot per cell, all distinct, below the parking slots
  std::set<long> cells;
  for (int j = 0; j < nJ; ++j)
    for (int i = 0; i < nI; ++i) {
      const long s = skewIndex(i, j, R);
      if (s < 0 || s >= parkStart) return fail("writer", nI, nJ, B, s, parkStart);
      if (!cells.insert(s).second) return fail("writer collision", nI, nJ, B, s, len);
    }
  for (int w = 0; w < W; ++w) {
    const int kBegin = 64 * w;
    const int kEnd = std::min(steps, 64 * w + 63 + nI);
    const int kFirst = kBegin - D;
    const int count = std::max(kEnd - kFirst, 0);
    const int rings = (count + D - 1) / D;
    for (int l = 0; l < 64; ++l) {
      const int lane = 64 * w + l;
      const int laneG = std::min(lane, R - 1);
      const bool active = lane < nJ;
      // the running offsets, in slots (the kernel keeps them in bytes, x16)
      long oF = static_cast<long>(kFirst + D + kSkewFront) * R + laneG;
      long oS = lane < R ? static_cast<long>(kFirst + kSkewFront) * R + lane : skewParkIndex(nI, nJ, R, lane);
      const long incS = lane < R ? R : 0;
      if (kFirst + kSkewFront < 0) return fail("front padding", nI, nJ, B, kFirst, 0);
      for (int k = kFirst; k < kFirst + rings * D; ++k) {
        const int i = k - lane;
        const bool on = active && i >= 0 && i < nI;
        // the slot this step uses was fetched D steps ago at (k + kSkewFront) R + laneG
        const long at = static_cast<long>(k + kSkewFront) * R + laneG;
        if (on) {
          if (at != skewIndex(i, lane, R)) return fail("fetch cell", nI, nJ, B, at, skewIndex(i, lane, R));
          if (i + 1 < nI && at + R != skewIndex(i + 1, lane, R)) return fail("fetch east", nI, nJ, B, at + R, 0);
          if (lane + 1 < nJ && at + R + 1 != skewIndex(i, lane + 1, R)) return fail("fetch north", nI, nJ, B, at + R + 1, 0);
        }
        // store of this step
        if (oS < 0 || oS >= len) return fail("store", nI, nJ, B, oS, len);
        if (lane < R) {
          if (oS != at) return fail("store slot", nI, nJ, B, oS, at);
          if (on != (cells.count(oS) == 1)) return fail("store off the wavefront hits a cell", nI, nJ, B, oS, 0);
        } else if (oS < parkStart) {
          return fail("parking", nI, nJ, B, oS, parkStart);
        }
        oS += incS;
        // fetch for step k + D: the slot, the next row, the next row's next slot
        const long oE = oF + R;
        if (oF < 0 || oE + 1 >= parkStart) return fail("fetch bounds", nI, nJ, B, oE + 1, parkStart);
        oF += R;
        // LDS: row 0 write (lane 0) and read-ahead, hand-off ring, partner map
        if (lane == 0 && on && i >= nI) return fail("lds row0", nI, nJ, B, i, nI);
        const int pIdx = std::min(std::max(i + 2, 0), nI - 1);
        if (pIdx < 0 || pIdx >= nI) return fail("lds partner", nI, nJ, B, pIdx, nI);
        const long handW = static_cast<long>(w) * 64 + (k & 63);
        if (handW < 0 || handW >= 64L * W) return fail("lds hand write", nI, nJ, B, handW, 64L * W);
        if (w > 0) {
          const long handR = static_cast<long>(w - 1) * 64 + (k & 63);
          if (handR < 0 || handR >= 64L * W) return fail("lds hand read", nI, nJ, B, handR, 64L * W);
        }
      }
    }
  }
  return 0;
}

// The colour-packed layout of the mixed path's two-colour sweeps (rbIndex):
// a bijection of the mesh's cells onto [0, N), colour 0 first.
static int checkColours(int nI, int nJ) {
  const int N = nI * nJ;
  const int split = rbRowBase(nJ, 0, nI);
  std::set<int> seen;
  for (int j = 0; j < nJ; ++j)
    for (int i = 0; i < nI; ++i) {
      const int k = rbIndex(i, j, nI, nJ);
      if (k < 0 || k >= N) return fail("colour index", nI, nJ, 0, k, N);
      if (((i + j) & 1) == 0 ? k >= split : k < split) return fail("colour plane", nI, nJ, 0, k, split);
      if (!seen.insert(k).second) return fail("colour index repeated", nI, nJ, 0, k, N);
    }
  return 0;
}

int main() {
  for (int nI = 1; nI <= 41; ++nI)
    for (int nJ = 1; nJ <= 41; ++nJ)
      if (checkColours(nI, nJ)) return 1;
  if (checkColours(1024, 256) || checkColours(512, 128) || checkColours(256, 64) || checkColours(118, 27)) return 1;
  const int shapes[][2] = {{1, 1}, {2, 1}, {1, 5}, {26, 5}, {26, 79}, {118, 27}, {64, 64}, {65, 63},
                           {256, 64}, {512, 128}, {1024, 256}, {300, 257}, {40, 1000}, {7, 1024}};
  for (const auto& s : shapes) {
    const int nI = s[0], nJ = s[1];
    for (int B : {64, 128, 256, 512, 1024}) {
      const int D = B <= 256 ? 8 : 3;  // k_gs_wave_* and k_gs_wave_*_wide
      if (check(nI, nJ, D, B)) return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
