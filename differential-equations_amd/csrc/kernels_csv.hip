// Solution::to_csv (src/solution.rs:190-230) for every trajectory of an
// ensemble solve, on the GPU.
//
// Work is cut into slots: slot r of trajectory i is its header (r = 0) or
// its point r - 1.  Three passes over a chunk of trajectories:
//   csv_points_kernel  the number of points of each trajectory and, for t_eval
//                      output, where they sit in t_eval;
//   csv_len_kernel     one thread per slot: the byte length of the slot's line
//                      (each number through fmt_prepare: Ryu digits, layout);
//   exclusive scan     of the lengths (three passes of the library's own:
//                      block sums, one block scanning them, block scans plus
//                      their offset; no pass waits on another block): every
//                      slot's output position, trajectory i's document
//                      starting at its header's;
//   csv_write_kernel   256 consecutive slots per block: each thread formats
//                      its line into LDS at its position, then the block copies
//                      its contiguous byte range out with aligned 4-byte stores
//                      (byte stores only at the range's ragged ends).
// The numbers are formatted twice (length, then text): a double's digits cost
// less to recompute than to store and reload.  Every loop is bounded (the
// digit search has fixed caps, see de_fmt.h); a formatter that reaches a cap
// sets a flag that the host reports as an error.
#include <hip/hip_runtime.h>

#include "de_fmt.h"
#include "de_launch.h"

namespace de {
namespace {

constexpr int kCsvBlock = 256;
// Slots per chunk: bounds the temporary arrays (12 bytes per slot).
constexpr int64_t kCsvChunkSlots = int64_t(1) << 26;

DE_DEV int header_len(int dim) {
  int n = 2;  // "t" and "\n"
  for (int c = 0; c < dim; ++c) n += 2 + (c >= 10 ? 2 : 1);
  return n;
}

DE_DEV int header_emit(int dim, char* dst) {
  int n = 0;
  dst[n++] = 't';
  for (int c = 0; c < dim; ++c) {
    dst[n++] = ',';
    dst[n++] = 'y';
    if (c >= 10) dst[n++] = (char)('0' + c / 10);
    dst[n++] = (char)('0' + c % 10);
  }
  dst[n++] = '\n';
  return n;
}

// Point p of trajectory i: its time and state.
DE_DEV void csv_point(const CsvArgs& a, int64_t i, int32_t p, int32_t head, int32_t run, double* t,
                      double* y) {
  const int64_t n = a.n;
  const int dim = a.dim;
  if (a.source == kCsvFinal) {
    *t = a.t_final[i];
    for (int c = 0; c < dim; ++c) y[c] = a.y_final[c * n + i];
  } else if (a.source == kCsvEvents) {
    *t = a.event_t[(int64_t)p * n + i];
    for (int c = 0; c < dim; ++c) y[c] = a.event_y[((int64_t)p * dim + c) * n + i];
  } else if (a.source == kCsvPoints) {
    const int64_t s = a.pts_offset[i] + p;
    *t = a.pts_t[s];
    for (int c = 0; c < dim; ++c) y[c] = a.pts_y[s * dim + c];
  } else {
    const int64_t k = p < head ? 0 : (int64_t)run + p - head;
    *t = a.t_eval[k];
    for (int c = 0; c < dim; ++c) y[c] = a.dense[(k * dim + c) * n + i];
  }
}

// The points of each trajectory of the chunk.  t_eval output: TEvalSolout's
// first call (t_prev = t_curr = t0) produces index 0 if it equals t0 and
// passes over every other point not beyond t0; every later point is produced
// in order (src/solout/t_eval.rs:100-135).  So the points are index 0 when
// `head`, then the run from index `run`.
__global__ void csv_points_kernel(CsvArgs a, int64_t i0, int64_t cnt, int32_t* count,
                                  int32_t* head, int32_t* run) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const int64_t i = i0 + j;
  int32_t c = 0, h = 0, r = 0;
  if (a.source == kCsvFinal) {
    c = 1;
  } else if (a.source == kCsvEvents) {
    c = a.n_events ? a.n_events[i] : 0;
  } else if (a.source == kCsvPoints) {
    const int64_t kept = a.pts_offset[i + 1] - a.pts_offset[i];
    const int64_t pushed = a.n_pts[i];
    c = (int32_t)(pushed < kept ? pushed : kept);
  } else {
    const double t0 = a.t0 ? a.t0[i] : a.t0s;
    const double tf = a.tf ? a.tf[i] : a.tfs;
    const bool fwd = tf > t0;
    h = a.n_eval > 0 && a.t_eval[0] == t0;
    int32_t idx = 0;
    while (idx < a.n_eval) {
      const double te = a.t_eval[idx];
      if ((fwd && te > t0) || (!fwd && te < t0)) break;
      ++idx;
    }
    r = idx;
    c = a.n_dense ? a.n_dense[i] : 0;
  }
  // Only the stored points: max_events of them, and the t_eval rows that exist.
  const int32_t cap = a.source == kCsvTeval ? h + (a.n_eval - r) : a.max_rows;
  count[j] = c < 0 ? 0 : (c > cap ? cap : c);
  head[j] = h;
  run[j] = r;
}

// Byte length of slot s of the chunk (0 past a trajectory's points);
// len[nslots] = 0 so that the scan's last entry is the chunk's total.
__global__ void csv_len_kernel(CsvArgs a, int64_t i0, uint32_t nslots, uint32_t per,
                               const int32_t* count, const int32_t* head, const int32_t* run,
                               int64_t* len, uint32_t* bad) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslots) return;
  if (s == nslots) {
    len[s] = 0;
    return;
  }
  const uint32_t j = s / per, r = s - j * per;
  const int32_t cnt = count[j];
  int32_t bytes = 0;
  if (cnt > 0) {
    if (r == 0) {
      bytes = header_len(a.dim);
    } else if ((int32_t)r - 1 < cnt) {
      double t, y[kCsvMaxDim];
      csv_point(a, i0 + j, (int32_t)r - 1, head[j], run[j], &t, y);
      FmtF64 f = fmt_prepare(t);
      int flag = f.bad;
      bytes = f.len + a.dim + 1;
      for (int c = 0; c < a.dim; ++c) {
        f = fmt_prepare(y[c]);
        flag |= f.bad;
        bytes += f.len;
      }
      if (flag) atomicOr(bad, 1u);
    }
  }
  len[s] = bytes;
}

// ---- exclusive scan of len[0 .. n) into pos[0 .. n) -----------------------
// kScanItems consecutive elements per thread, kCsvBlock threads per block.
constexpr int kScanItems = 8;
constexpr int64_t kScanTile = (int64_t)kScanItems * kCsvBlock;

// Exclusive scan of the block's kCsvBlock values in LDS; returns the total.
DE_DEV int64_t block_exclusive_scan(int64_t* sh, int64_t v) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < kCsvBlock; off <<= 1) {  // inclusive Hillis-Steele
    const int64_t add = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  const int64_t total = sh[kCsvBlock - 1];
  const int64_t excl = sh[tid] - v;
  __syncthreads();
  sh[tid] = excl;
  return total;
}

// Pass 1: sums[b] = sum of tile b.
__global__ __launch_bounds__(kCsvBlock) void scan_tile_sums(const int64_t* len, int64_t n,
                                                            int64_t* sums) {
  __shared__ int64_t sh[kCsvBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) v += len[base + k];
  const int64_t total = block_exclusive_scan(sh, v);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Pass 2, one block: sums[0 .. tiles) replaced by their exclusive scan.
__global__ __launch_bounds__(kCsvBlock) void scan_tile_offsets(int64_t* sums, int64_t tiles) {
  __shared__ int64_t sh[kCsvBlock];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < tiles; b0 += kScanTile) {  // tiles <= 2^16: a few rounds
    const int64_t base = b0 + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems], tsum = 0;
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < tiles ? sums[base + k] : 0;
      tsum += v[k];
    }
    __syncthreads();
    const int64_t total = block_exclusive_scan(sh, tsum);
    int64_t run = carry + sh[threadIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < tiles) sums[base + k] = run;
      run += v[k];
    }
    carry += total;
  }
}

// Pass 3: pos = the tile's exclusive scan plus the tile's offset.
__global__ __launch_bounds__(kCsvBlock) void scan_tile_apply(const int64_t* len, int64_t n,
                                                             const int64_t* offs, int64_t* pos) {
  __shared__ int64_t sh[kCsvBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems], tsum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? len[base + k] : 0;
    tsum += v[k];
  }
  block_exclusive_scan(sh, tsum);
  int64_t run = offs[blockIdx.x] + sh[threadIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) pos[base + k] = run;
    run += v[k];
  }
}

// offsets[i0 + j] = base + pos[j * per]; the chunk's total is added to base.
__global__ void csv_offsets_kernel(int64_t cnt, uint32_t per, uint32_t nslots, const int64_t* pos,
                                   int64_t* base, int64_t* offsets, int64_t i0, int last) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = *base;
  if (j < cnt) offsets[i0 + j] = b + pos[j * per];
  if (j == cnt && last) offsets[i0 + cnt] = b + pos[nslots];
}

__global__ void csv_base_kernel(uint32_t nslots, const int64_t* pos, int64_t* base) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *base += pos[nslots];
}

__global__ void csv_zero_kernel(int64_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

__global__ __launch_bounds__(kCsvBlock) void csv_write_kernel(
    CsvArgs a, int64_t i0, uint32_t nslots, uint32_t per, const int32_t* count,
    const int32_t* head, const int32_t* run, const int64_t* pos, const int64_t* base, char* out,
    int64_t cap, uint32_t* bad) {
  extern __shared__ char lds[];
  const uint32_t s0 = blockIdx.x * kCsvBlock;
  const uint32_t s = s0 + threadIdx.x;
  const uint32_t s1 = s0 + kCsvBlock < nslots ? s0 + kCsvBlock : nslots;
  const int64_t b = *base;
  // A block whose bytes would pass the caller's buffer writes nothing (the
  // host then reports the size mismatch).
  if (b + pos[s1] > cap) return;
  // The block's bytes [g_lo, g_hi) of out, staged in LDS from the 4-byte
  // aligned address at or below g_lo.
  const uintptr_t g_lo = (uintptr_t)(out + b + pos[s0]);
  const uintptr_t g_hi = (uintptr_t)(out + b + pos[s1]);
  const uintptr_t g_al = g_lo & ~(uintptr_t)3;
  if (s < nslots) {
    const int64_t p0 = pos[s], bytes = pos[s + 1] - p0;
    if (bytes > 0) {
      // Texts are formed in a private buffer and copied to LDS one byte at a
      // time through a volatile pointer: the slot starts at any byte, and
      // consecutive byte stores that the compiler merged into wider LDS stores
      // at unaligned addresses lost bytes.
      volatile char* dst = lds + ((uintptr_t)(out + b + p0) - g_al);
      char buf[ryu::kMaxChars + 8];
      const uint32_t j = s / per, r = s - j * per;
      if (r == 0) {
        const int hl = header_emit(a.dim, buf);
        for (int q = 0; q < hl && q < bytes; ++q) dst[q] = buf[q];
        if (hl != bytes) atomicOr(bad, 1u);
      } else {
        double t, y[kCsvMaxDim];
        csv_point(a, i0 + j, (int32_t)r - 1, head[j], run[j], &t, y);
        int flag = 0, k = 0;
        for (int c = -1; c < a.dim; ++c) {
          if (c >= 0) {
            if (k < bytes) dst[k] = ',';
            ++k;
          }
          const FmtF64 f = fmt_prepare(c < 0 ? t : y[c]);
          flag |= f.bad;
          fmt_emit(f, buf);
          for (int q = 0; q < f.len; ++q, ++k)
            if (k < bytes) dst[k] = buf[q];
        }
        if (k < bytes) dst[k] = '\n';
        // The length pass measured the same text: a differing total is a bug,
        // reported like a cap (and no byte past the slot is written).
        if (flag || k + 1 != bytes) atomicOr(bad, 1u);
      }
    }
  }
  __syncthreads();
  const uintptr_t w_lo = (g_lo + 3) & ~(uintptr_t)3;  // first whole word
  const uintptr_t w_hi = g_hi & ~(uintptr_t)3;        // end of the whole words
  if (w_lo >= w_hi) {
    for (uintptr_t g = g_lo + threadIdx.x; g < g_hi; g += kCsvBlock)
      *(char*)g = lds[g - g_al];
    return;
  }
  if (threadIdx.x < w_lo - g_lo) *(char*)(g_lo + threadIdx.x) = lds[g_lo + threadIdx.x - g_al];
  if (threadIdx.x < g_hi - w_hi) *(char*)(w_hi + threadIdx.x) = lds[w_hi + threadIdx.x - g_al];
  for (uintptr_t g = w_lo + 4 * (uintptr_t)threadIdx.x; g < w_hi; g += 4 * kCsvBlock)
    *(uint32_t*)g = *(const uint32_t*)(lds + (g - g_al));
}

// One text per value into a 32-byte slot; len[i] = its length, negated when
// the digit search reached a cap.
__global__ void diag_fmt_kernel(const double* x, char* out, int32_t* len, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FmtF64 f = fmt_prepare(x[i]);
  char buf[32];
  fmt_emit(f, buf);
  for (int k = 0; k < f.len && k < 32; ++k) out[i * 32 + k] = buf[k];
  len[i] = f.bad ? -f.len : f.len;
}

}  // namespace

int launch_fmt_diag(const double* x, char* out, int32_t* len, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(diag_fmt_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, out,
                     len, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Sizes (offsets) and, when out is not null, the documents of every
// trajectory into out[0 .. cap); device pointers.  Returns 0, -1 on a HIP
// error, -2 when a formatter reached a loop cap.
int launch_csv(const CsvArgs& a, int64_t* offsets, char* out, int64_t cap, hipStream_t s) {
  if (a.n == 0) {
    hipLaunchKernelGGL(csv_zero_kernel, dim3(1), dim3(1), 0, s, offsets);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  const int64_t per = (int64_t)a.max_rows + 1;
  int64_t chunk = kCsvChunkSlots / per;
  if (chunk < 1) chunk = 1;
  if (chunk > a.n) chunk = a.n;
  const int64_t max_slots = chunk * per;
  const int64_t max_tiles = (max_slots + 1 + kScanTile - 1) / kScanTile;
  // One workspace: count/head/run, len, pos, the tile sums, the running base
  // and the formatter's flag.
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_chr = up(sizeof(int32_t) * 3 * (size_t)chunk);
  const size_t b_len = up(sizeof(int64_t) * (size_t)(max_slots + 1));
  const size_t b_sum = up(sizeof(int64_t) * (size_t)max_tiles);
  char* w = nullptr;
  hipError_t e = workspace_alloc((void**)&w, b_chr + 2 * b_len + b_sum + 256, s);
  if (e != hipSuccess) return -1;
  int32_t* cnt_head_run = (int32_t*)w;
  int64_t* len = (int64_t*)(w + b_chr);
  int64_t* pos = (int64_t*)(w + b_chr + b_len);
  int64_t* sums = (int64_t*)(w + b_chr + 2 * b_len);
  int64_t* base = (int64_t*)(w + b_chr + 2 * b_len + b_sum);
  uint32_t* bad = (uint32_t*)(base + 1);
  hipLaunchKernelGGL(csv_zero_kernel, dim3(1), dim3(1), 0, s, base);
  hipLaunchKernelGGL(csv_zero_kernel, dim3(1), dim3(1), 0, s, (int64_t*)bad);
  const size_t lds = (size_t)kCsvBlock * csv_max_slot_bytes(a.dim) + 8;
  e = hipGetLastError();
  for (int64_t i0 = 0; e == hipSuccess && i0 < a.n; i0 += chunk) {
    const int64_t cnt = a.n - i0 < chunk ? a.n - i0 : chunk;
    const uint32_t nslots = (uint32_t)(cnt * per);
    const int64_t nscan = (int64_t)nslots + 1;
    const int64_t tiles = (nscan + kScanTile - 1) / kScanTile;
    int32_t* count = cnt_head_run;
    int32_t* head = cnt_head_run + chunk;
    int32_t* run = cnt_head_run + 2 * chunk;
    hipLaunchKernelGGL(csv_points_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, a,
                       i0, cnt, count, head, run);
    hipLaunchKernelGGL(csv_len_kernel, dim3((nslots + 1 + 255) / 256), dim3(256), 0, s, a, i0,
                       nslots, (uint32_t)per, (const int32_t*)count, (const int32_t*)head,
                       (const int32_t*)run, len, bad);
    hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)tiles), dim3(kCsvBlock), 0, s,
                       (const int64_t*)len, nscan, sums);
    hipLaunchKernelGGL(scan_tile_offsets, dim3(1), dim3(kCsvBlock), 0, s, sums, tiles);
    hipLaunchKernelGGL(scan_tile_apply, dim3((unsigned)tiles), dim3(kCsvBlock), 0, s,
                       (const int64_t*)len, nscan, (const int64_t*)sums, pos);
    const int last = i0 + cnt == a.n;
    hipLaunchKernelGGL(csv_offsets_kernel, dim3((unsigned)((cnt + 1 + 255) / 256)), dim3(256), 0,
                       s, cnt, (uint32_t)per, nslots, (const int64_t*)pos, base, offsets, i0, last);
    if (out)
      hipLaunchKernelGGL(csv_write_kernel, dim3((nslots + kCsvBlock - 1) / kCsvBlock),
                         dim3(kCsvBlock), lds, s, a, i0, nslots, (uint32_t)per,
                         (const int32_t*)count, (const int32_t*)head, (const int32_t*)run,
                         (const int64_t*)pos, (const int64_t*)base, out, cap, bad);
    hipLaunchKernelGGL(csv_base_kernel, dim3(1), dim3(1), 0, s, nslots, (const int64_t*)pos, base);
    e = hipGetLastError();
  }
  uint32_t flag = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&flag, bad, sizeof flag, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  const hipError_t f = hipFreeAsync(w, s);
  if (e == hipSuccess) e = f;
  if (e != hipSuccess) return -1;
  return flag ? -2 : 0;
}

}  // namespace de
