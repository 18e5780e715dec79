// Jonker-Volgenant fast-elimination edge test on gfx950.
//
// Reference kernel: JvCandidatesKernel (src/cuda/jv_cuda.cu:221-288), one
// thread per edge slot walking N(a), N(b) and every candidate's row serially.
//
// Here a 64-lane wavefront owns two edge slots, one per 32-lane half, when
// both rows of each fit its half (deg_a + deg_b <= 32, the usual case for
// k-NN graphs), and one edge slot otherwise.  A half whose edge turns out
// not to fit (a longer row, a key that does not pack into 64 bits, a
// candidate row longer than 32) leaves that edge to a full wavefront after
// the pair, which repeats the edge from the start on the one-edge path.
//
// One edge slot (a, b) on a full wavefront:
//   1. candidate scoring (jv_elimination.cpp:45-84): lane s scores the s-th
//      entry of N(a) followed by N(b); x in {a, b} is skipped, and on the b
//      side x is skipped when (a, x) is an edge (the reference's HasActiveEdge
//      binary search; here a broadcast scan of N(a) staged in LDS when both
//      rows fit one pass).  score = w(side, x) + D(x, other).  Each lane also
//      fetches x's row bounds and D(a, x), D(b, x), which the witness test
//      needs if x becomes a candidate.
//   2. top ten by (score, node): when both rows fit one pass and every key
//      packs into 64 bits, each valid lane counts the smaller keys with
//      broadcast LDS reads; otherwise by readlane.  Ranks < 10 land in LDS in
//      ascending order with their row bounds and distances.  Rows longer than
//      one pass merge into the running list.
//   3. witness test (jv_elimination.cpp:86-117): candidates in rank order,
//      four per pass in 16-lane slots when every candidate row has at most 16
//      entries, else as many consecutive candidates per pass as their rows
//      fit 64 lanes; lane j tests one neighbour d of its candidate c and a
//      ballot tells,
//      per candidate, whether any d is compatible (w_ab + w_cd <= D(a,c) +
//      D(d,b) or w_ab + w_cd <= D(a,d) + D(b,c)).  The first c with no
//      compatible d is the witness.  Testing every d at once gives the same
//      verdict as the reference's early exit at the first compatible d, and
//      testing later candidates of a pass does not change which is first.
//      (The pair path below first probes three neighbours of every candidate
//      at once, then finishes the rows of the candidates still open in rank
//      order.)
// The pair path runs the same three steps in each half: scoring by lane
// (N(a) staged in the half's LDS row), ranking by broadcast LDS reads of the
// half's keys, and the witness test as a probe pass (three lanes per
// candidate, the first three entries of its row) followed by one pass per
// candidate still open, lowest rank first, over the rest of its row.
// Sums stay exact in uint64: the host guarantees weights <= 2^62 and
// distances <= 2^32, so no sum wraps and unsigned comparison equals the
// reference CPU's 128-bit comparison.
#include "ee_jv_body.h"

namespace ee {
namespace {

using namespace jvbody;

// Two edge slots per wavefront (edges that do not fit a half run afterwards
// on the whole wavefront), or one per wavefront when g.one_per_wave is set.
// Six wavefronts per SIMD (at most 80 VGPRs): the kernel is bound by its
// chains of dependent gathers, and more wavefronts in flight hide more of
// them (lattice epoch 24.3 -> 22.6 us against the compiler's own choice; with
// the probe witness pass, 55 us against 58 us at five per SIMD on a
// 15,000-node graph, 19.5 against 18.9 us on the lattice).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) JvKernel(JvLaunch g) {
  __shared__ JvShared sh;
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const int wib = static_cast<int>(threadIdx.x) / kWave;
  const int32_t waves = static_cast<int32_t>(gridDim.x) * kWavesPerBlock;
  const int32_t first = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock + wib);
  if (g.one_per_wave == 0) {
    const int32_t pairs = (g.edge_count + 1) / 2;
    for (int32_t pair = first; pair < pairs; pair += waves) {
      const uint64_t rest = __ballot(PairOfEdges(g, sh, wib, lane, pair));
      for (int h = 0; h < 2; ++h) {  // uniform
        if (((rest >> (kHalf * h)) & 1ull) != 0) {
          const int32_t item = 2 * pair + h;
          const int32_t result = OneEdge(g, sh, wib, lane, item);
          if (lane == 0) g.witnesses[item] = result;
        }
      }
    }
    return;
  }
  for (int32_t item = first; item < g.edge_count; item += waves) {
    const int32_t result = OneEdge(g, sh, wib, lane, item);
    if (lane == 0) g.witnesses[item] = result;
  }
}

}  // namespace

void LaunchJv(const JvLaunch& args, uint32_t grid, hipStream_t stream) {
  if (args.edge_count <= 0) return;
  // one wavefront per edge slot, or per pair of edge slots
  const uint64_t items = args.one_per_wave != 0 ? static_cast<uint64_t>(args.edge_count)
                                                : (static_cast<uint64_t>(args.edge_count) + 1) / 2;
  const uint64_t needed = (items + kWavesPerBlock - 1) / kWavesPerBlock;
  uint64_t blocks = grid == 0 ? needed : grid;
  if (blocks > needed) blocks = needed;
  hipLaunchKernelGGL(JvKernel, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, stream, args);
  CheckHip(hipGetLastError(), "JV kernel launch");
}

}  // namespace ee
