// Many materials, one launch: a persistent kernel drains a queue of tiles
// (material, range of that material's neutrons).  Each material keeps its
// own composition view, RNG seed and call counter, and each of its neutrons
// draws from the same Philox stream as in a per-material call, so a batched
// launch returns exactly what one call per material returns.
//
// Tiles are handed out by one atomic counter in the order of a tile list
// built on the host (material order, or most expensive materials first so
// that the tail of the launch is made of short tiles); the tile size, the
// number of resident workgroups and the order are the autotuned parameters.
#include "nc_kernels.hh"

namespace ncamd {

// Cross sections: one persistent kernel.  Sampling in three stages so that
// no kernel carries every sampler's registers (the S(alpha,beta) sampler
// alone would hold the occupancy at one wave per SIMD):
//   k_batch_pick  picks the component, samples PowderBragg / ElInc in place
//                 and lists each tile's S(alpha,beta) neutrons (block-wide
//                 compaction, so each list is in neutron order);
//   k_batch_sab   samples the listed neutrons tile by tile (one material per
//                 workgroup), leaving those above the table's Emax in a queue;
//   k_batch_highe samples that queue with the free-gas extender.
// The S(alpha,beta) stages restart the neutron's Philox stream and replay the
// pick's draw (one draw when the material has several components), so every
// neutron gets what a single kernel would give it.
constexpr int kBatchBlock = 256;

__device__ __forceinline__ int batch_next_tile(unsigned* next, int* s_tile) {
  if (threadIdx.x == 0) *s_tile = (int)atomicAdd(next, 1u);
  __syncthreads();
  const int t = *s_tile;
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(kBatchBlock) k_batch_xs(const BatchMat* __restrict__ mats,
                                                          const BatchTile* __restrict__ tiles, int ntiles,
                                                          const double* __restrict__ ekin, double* __restrict__ o1,
                                                          unsigned* __restrict__ next) {
  __shared__ int s_tile;
  while (true) {
    const int t = batch_next_tile(next, &s_tile);
    if (t >= ntiles) break;  // every workgroup leaves once the list is drained
    const BatchTile tile = tiles[t];
    const BatchMat& m = mats[tile.mat];
    for (int64_t k = tile.begin + threadIdx.x; k < tile.end; k += blockDim.x)
      o1[m.offset + k] = comp_xs<false>(m.view, ekin[m.offset + k], nullptr);
  }
}

__global__ void __launch_bounds__(kBatchBlock) k_batch_pick(const BatchMat* __restrict__ mats,
                                                            const BatchRng* __restrict__ rngs,
                                                            const BatchTile* __restrict__ tiles, int ntiles,
                                                            const double* __restrict__ ekin, double* __restrict__ o1,
                                                            double* __restrict__ o2, unsigned* __restrict__ next,
                                                            int64_t* __restrict__ sab, int* __restrict__ tcount,
                                                            unsigned long long* __restrict__ nerr) {
  __shared__ int s_tile;
  __shared__ int s_wave[kBatchBlock / 64];
  const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
  const uint64_t below = (1ull << lane) - 1ull;
  while (true) {
    const int t = batch_next_tile(next, &s_tile);
    if (t >= ntiles) break;
    const BatchTile tile = tiles[t];
    const BatchMat& m = mats[tile.mat];
    const BatchRng r = rngs[tile.mat];
    int run = 0;  // S(alpha,beta) neutrons listed so far in this tile
    for (int64_t start = tile.begin; start < tile.end; start += blockDim.x) {
      const int64_t k = start + threadIdx.x;
      bool to_sab = false;
      if (k < tile.end) {
        Rng rng;
        rng.init(r.seed, (uint32_t)r.stream, (uint64_t)k);
        double Ef, mu;
        int slot = 0;  // a batch material has at most one staged component, its S(alpha,beta) one
        const int rc = comp_pick_sample<false>(m.view, rng, ekin[m.offset + k], Ef, mu, slot);
        if (rc == kCompToStage) {
          to_sab = true;
        } else {
          if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
          o1[m.offset + k] = Ef;
          o2[m.offset + k] = mu;
        }
      }
      const uint64_t bal = __ballot(to_sab);
      if (lane == 0) s_wave[wid] = __popcll(bal);
      __syncthreads();
      int off = run, sum = 0;
      for (int w = 0; w < kBatchBlock / 64; ++w) {
        if (w < wid) off += s_wave[w];
        sum += s_wave[w];
      }
      if (to_sab) sab[m.offset + tile.begin + off + __popcll(bal & below)] = k;
      __syncthreads();  // s_wave is rewritten by the next chunk
      run += sum;
    }
    if (threadIdx.x == 0) tcount[t] = run;
  }
}

__global__ void __launch_bounds__(kBatchBlock) k_batch_sab(const BatchMat* __restrict__ mats,
                                                           const BatchRng* __restrict__ rngs,
                                                           const BatchTile* __restrict__ tiles, int ntiles,
                                                           const double* __restrict__ ekin, double* __restrict__ o1,
                                                           double* __restrict__ o2, unsigned* __restrict__ next,
                                                           const int64_t* __restrict__ sab,
                                                           const int* __restrict__ tcount,
                                                           int64_t* __restrict__ highe,
                                                           unsigned long long* __restrict__ hcount,
                                                           unsigned long long* __restrict__ nerr) {
  __shared__ int s_tile;
  while (true) {
    const int t = batch_next_tile(next, &s_tile);
    if (t >= ntiles) break;
    const BatchTile tile = tiles[t];
    const BatchMat& m = mats[tile.mat];
    const int cnt = tcount[t];
    if (cnt == 0) continue;
    const BatchRng r = rngs[tile.mat];
    const int skip = m.view.n > 1 ? 1 : 0;
    const double emax = m.view.sab[0].egrid[m.view.sab[0].ne - 1];
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
      const int64_t k = sab[m.offset + tile.begin + j];
      const double E = ekin[m.offset + k];
      if (!(E < emax)) {
        highe[atomicAdd(hcount, 1ull)] = ((int64_t)t << 40) | k;
        continue;
      }
      Rng rng;
      rng.init(r.seed, (uint32_t)r.stream, (uint64_t)k);
      if (skip) rng.gen();
      double Ef, mu;
      const int rc = sab_sample_iso<false>(m.view.sab[0], rng, E, Ef, mu);
      if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
      o1[m.offset + k] = Ef;
      o2[m.offset + k] = mu;
    }
  }
}

__global__ void __launch_bounds__(kBatchBlock, 2) k_batch_highe(const BatchMat* __restrict__ mats,
                                                             const BatchRng* __restrict__ rngs,
                                                             const BatchTile* __restrict__ tiles,
                                                             const double* __restrict__ ekin, double* __restrict__ o1,
                                                             double* __restrict__ o2,
                                                             const int64_t* __restrict__ highe,
                                                             const unsigned long long* __restrict__ hcount,
                                                             unsigned long long* __restrict__ nerr) {
  const int64_t nq = (int64_t)*hcount;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t item = highe[q];
    const int t = (int)(item >> 40);
    const int64_t k = item & ((1ll << 40) - 1);
    const BatchTile tile = tiles[t];
    const BatchMat& m = mats[tile.mat];
    const BatchRng r = rngs[tile.mat];
    Rng rng;
    rng.init(r.seed, (uint32_t)r.stream, (uint64_t)k);
    if (m.view.n > 1) rng.gen();
    double Ef, mu;
    const int rc = sab_sample_iso<true>(m.view.sab[0], rng, ekin[m.offset + k], Ef, mu);
    if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
    o1[m.offset + k] = Ef;
    o2[m.offset + k] = mu;
  }
}

__global__ void k_batch_reset(unsigned* next, unsigned long long* hcount) {
  next[0] = 0u;
  next[1] = 0u;
  *hcount = 0ull;
}

void launch_batch(const BatchMat* mats, const BatchRng* rngs, const BatchTile* tiles, int ntiles, int nblocks,
                  int mode, const double* ekin, double* o1, double* o2, const BatchWork& w,
                  unsigned long long* nerr, hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, st, w.next, w.hcount);
  if (mode == 0) {
    hipLaunchKernelGGL(k_batch_xs, dim3(nblocks), dim3(kBatchBlock), 0, st, mats, tiles, ntiles, ekin, o1, w.next);
    return;
  }
  hipLaunchKernelGGL(k_batch_pick, dim3(nblocks), dim3(kBatchBlock), 0, st, mats, rngs, tiles, ntiles, ekin, o1, o2,
                     w.next, w.sab, w.tcount, nerr);
  hipLaunchKernelGGL(k_batch_sab, dim3(nblocks), dim3(kBatchBlock), 0, st, mats, rngs, tiles, ntiles, ekin, o1, o2,
                     w.next + 1, w.sab, w.tcount, w.highe, w.hcount, nerr);
  hipLaunchKernelGGL(k_batch_highe, dim3(nblocks), dim3(kBatchBlock), 0, st, mats, rngs, tiles, ekin, o1, o2,
                     w.highe, w.hcount, nerr);
}

}  // namespace ncamd
