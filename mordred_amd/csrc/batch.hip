// Shared-scan batch kernels for gfx950 (see batch.hpp for the dimension
// tables and the batch descriptor).
//
//   batch_fill_kernel     clears the dimension tables of the used dimensions
//                         (when rebuilt), and the result rows and touched
//                         bitmaps of the queries being started;
//   batch_records_kernel  one pass over a dimension: per row its slot, the
//                         mask of the queries whose filter it passes, and its
//                         packed group attributes (the builds of all the
//                         batch's queries in one);
//   batch_kernel          the persistent shared scan of the fact columns;
//   batch_keys_kernel     the keys of every touched group, decoded from the
//                         group index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "batch.hpp"
#include "pipeline.hpp"
#include "tile.hpp"

namespace mordred {

// Loads through a global (address space 1) pointer: the table pointers come
// from LDS-staged descriptors, so the compiler would otherwise emit flat
// loads, which wait for every earlier LDS and memory access before issuing.
template <typename T>
__device__ __forceinline__ T gload(const void* p, long long i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}

// A wave-uniform pointer held in scalar registers, so that loads from it take
// a 32-bit per-lane byte offset (one VGPR per address instead of two).
__device__ __forceinline__ const void* uniform_ptr(const void* p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return reinterpret_cast<const void*>(((unsigned long long)hi << 32) | lo);
}

// Global load at a 32-bit byte offset from a uniform base.
template <typename T>
__device__ __forceinline__ T gload_at(const void* base, uint32_t byte_off) {
  return *(const __attribute__((address_space(1))) T*)((const __attribute__((address_space(1))) char*)base + byte_off);
}

// Attribute-table entry of slot h (pw bytes).
__device__ __forceinline__ uint32_t gload_pw(const void* pk, int pw, int h) {
  return pw == 1 ? gload<uint8_t>(pk, h) : pw == 2 ? gload<uint16_t>(pk, h) : gload<uint32_t>(pk, h);
}

// Clears the dimension tables of the used dimensions when they are rebuilt,
// and the result rows and touched bitmaps of the queries being (re)started.
__global__ __launch_bounds__(256) void batch_fill_kernel(const BatchDesc* __restrict__ D, unsigned* status) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) status[0] = 0;
  const uint4 z = make_uint4(0, 0, 0, 0);
  if (D->rebuild) {
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const BatchDim& B = D->dim[d];
      if (!B.used) continue;
      uint4* mk = reinterpret_cast<uint4*>(B.mk);
      const long long nm = B.mstride == 2 ? (long long)B.len + 1 : ((long long)B.len + 2) / 2;  // slots [0, len]
      if (mk)
        for (long long i = t; i < nm; i += stride) mk[i] = z;
      uint4* pk = reinterpret_cast<uint4*>(B.pk);
      if (pk)
        for (long long i = t; i < B.pk_words; i += stride) pk[i] = z;
    }
  }
  const unsigned long long reset = D->reset;
  const int nq = D->nq;
  for (int q = 0; q < nq; q++) {
    if (!(reset >> q & 1ull)) continue;
    const BatchQuery& Q = D->q[q];
    int* res = reinterpret_cast<int*>(Q.res);
    const long long ni = 6 * (long long)Q.total_val;
    for (long long i = t; i < ni; i += stride) res[i] = 0;
    const int nt = batch_touch_words(Q.total_val, Q.flags);
    unsigned* tb = D->touch_arena + Q.gtouch;
    for (long long i = t; i < nt; i += stride) tb[i] = 0;
  }
}

// One thread per dimension row: its slot, the mask of the queries whose
// filter the row passes (build_GPU2's modes, GPUProcessing.h:972-980: 0 EQ,
// 1 BETWEEN, 2 EQ-or-EQ), and its packed group attributes.  The host has
// checked that no two rows share a slot.
__global__ __launch_bounds__(256) void batch_records_kernel(BatchLaunch L, int d) {
  const BatchDim& B = L.desc->dim[d];
  if (!B.used || !L.desc->rebuild || B.form == kDimDict) return;  // (dictionary codes: built by the host)
  const int nrows = B.nrows, len = B.len, mk = B.minkey, pw = B.pw, na = B.nattr, nfc = B.nfcol, nf = B.nfilt;
  const int stride = gridDim.x * blockDim.x;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < nrows; row += stride) {
    const int key = gather(L.cache, B.key, row);
    const int h = phash(key, len, mk);
    if ((unsigned)h >= (unsigned)len) continue;  // key below min: never probed in range
    if (B.form == kDimAttr) {
      // every attribute as value - base + 1: an occupied slot is never 0
      uint32_t w = na ? 0u : 1u;
      for (int a = 0; a < na; a++) {
        const BatchAttr A = B.attr[a];
        const uint32_t v = (uint32_t)(gather(L.cache, A.col, row) - A.base) + 1u;
        w |= (v & ((1u << A.bits) - 1u)) << A.off;
      }
      if (pw == 1) static_cast<uint8_t*>(B.pk)[h] = (uint8_t)w;
      else if (pw == 2) static_cast<uint16_t*>(B.pk)[h] = (uint16_t)w;
      else static_cast<uint32_t*>(B.pk)[h] = w;
      continue;
    }
    int fv[kBatchMaxFCol];
#pragma unroll
    for (int c = 0; c < kBatchMaxFCol; c++) fv[c] = c < nfc ? gather(L.cache, B.fcol[c], row) : 0;
    unsigned long long m = B.uses;
    for (int f = 0; f < nf; f++) {
      const BatchFilt F = B.filt[f];
      int x = key;
#pragma unroll
      for (int c = 0; c < kBatchMaxFCol; c++)
        if (F.src == c) x = fv[c];
      const bool ok = F.mode == 0 ? x == F.c1 : F.mode == 1 ? (x >= F.c1 && x <= F.c2) : (x == F.c1 || x == F.c2);
      if (!ok) m &= ~(1ull << F.q);
    }
    if (B.mstride == 1) {
      B.mk[h] = m;
      continue;
    }
    uint32_t w = 0;
    for (int a = 0; a < na; a++) {
      const BatchAttr A = B.attr[a];
      const uint32_t v = (uint32_t)(gather(L.cache, A.col, row) - A.base);
      w |= (v & ((1u << A.bits) - 1u)) << A.off;
    }
    reinterpret_cast<uint4*>(B.mk)[h] = make_uint4((unsigned)m, (unsigned)(m >> 32), w, 0u);
  }
}

// Group keys of every touched group of the queries whose keys decode from the
// group index: digit d of index h is (h / uniq[d]) % krange[d], the key word
// minv[d] + digit (0 for a dimension without group value).  blockIdx.y = query.
__global__ __launch_bounds__(256) void batch_keys_kernel(const BatchDesc* __restrict__ D) {
  const int q = blockIdx.y;
  if (q >= D->nq || !((D->all | D->aliases) >> q & 1ull)) return;
  const BatchQuery& Q = D->q[q];
  if (Q.flags & (kQueryAggr | kQueryKeysInLoop)) return;
  const BatchKeys K = D->keys[q];
  const unsigned* tb = D->touch_arena + Q.gtouch;
  int* res = reinterpret_cast<int*>(Q.res);
  const int G = Q.total_val;
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < G; h += gridDim.x * blockDim.x) {
    if (!(tb[h >> 5] >> (h & 31) & 1u)) continue;
    int kk[4];
#pragma unroll
    for (int d = 0; d < 4; d++) kk[d] = Q.uniq[d] != 0 ? K.minv[d] + (h / Q.uniq[d]) % K.krange[d] : 0;
    int2* key = reinterpret_cast<int2*>(res + 6 * (size_t)h);
    key[0] = make_int2(kk[0], kk[1]);
    key[1] = make_int2(kk[2], kk[3]);
  }
}

// The shared scan.  A workgroup walks one contiguous range of tiles.  The
// next tile's columns are loaded once the current tile's masks, group
// attributes and aggregate values are known (the columns are dead then), so
// the loads are in flight during the pair work.  A wave keeps the cache slots
// of the current segment's columns in one VGPR (lane c = column c) and
// reloads them only when the segment changes.
template <int BLOCK, int VPT, int MAXC>
__global__ __launch_bounds__(BLOCK) void batch_kernel(BatchLaunch L) {
  constexpr int IPT = 4 * VPT;
  constexpr int TILE = BLOCK * IPT;
  const BatchDesc* __restrict__ D = L.desc;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ BatchHot H;
  const int nq = D->nq;
  // query descriptors transposed into four arrays of 16-byte chunks (chunk c
  // of query q at lqc[c * kBatchMaxQ + q]): lanes of one wave read the
  // descriptors of different queries, which a 64-byte stride would put in the
  // same banks; the fixed stride makes the chunk offsets immediates
  // With uniq16 the chunks are repacked so that the pair step reads two:
  // chunk 0 = total_val, acc, hbase, expr | sink << 8 | touch << 16; chunk 1 =
  // shift bytes, bits bytes, uniq[0] | uniq[1] << 16, uniq[2] | uniq[3] << 16;
  // chunk 2 = res; chunk 3 as in BatchQuery.
  uint4* lqc = reinterpret_cast<uint4*>(lds_raw);
  const bool lq16 = D->uniq16 != 0;
  auto query = [&](int q, bool full) {
    union {
      BatchQuery Q;
      uint4 c[4];
    } u;
    if (lq16) {
      const uint4 a = lqc[q], b = lqc[kBatchMaxQ + q], r = lqc[2 * kBatchMaxQ + q];
      u.c[0] = make_uint4(r.x, r.y, a.x, a.y);
      u.c[1] = make_uint4(a.z, a.w, b.x, b.y);
      u.c[2] = make_uint4(b.z & 0xFFFFu, b.z >> 16, b.w & 0xFFFFu, b.w >> 16);
    } else {
      u.c[0] = lqc[q];
      u.c[1] = lqc[kBatchMaxQ + q];
      u.c[2] = lqc[2 * kBatchMaxQ + q];
    }
    u.c[3] = full ? lqc[3 * kBatchMaxQ + q] : make_uint4(0, 0, 0, 0);
    return u.Q;
  };
  // interval and mask tables, then the group region (per query a touched
  // bitmap, then LDS accumulators, 8-byte aligned; BatchQuery holds their
  // offsets from lds_raw), and at the end of the region the waves' pair lists
  unsigned* ltab = reinterpret_cast<unsigned*>(lds_raw + kBatchMaxQ * sizeof(BatchQuery));
  const int ntab = (D->ntab + 1) & ~1;
  unsigned* lgrp = ltab + ntab;
  const int ngrp = D->lds_words;
  {
    const uint4* src = reinterpret_cast<const uint4*>(D->q);
    for (int i = threadIdx.x; i < nq * 4; i += BLOCK) {
      const int c = i & 3;
      uint4 v = src[i];
      if (lq16 && c < 3) {
        const uint4 w0 = src[i - c], w1 = src[i - c + 1], w2 = src[i - c + 2];
        v = c == 0   ? make_uint4(w0.z, w0.w, w1.x, w1.y)
            : c == 1 ? make_uint4(w1.z, w1.w, w2.x | w2.y << 16, w2.z | w2.w << 16)
                     : make_uint4(w0.x, w0.y, 0u, 0u);
      }
      lqc[c * kBatchMaxQ + (i >> 2)] = v;
    }
    for (int i = threadIdx.x; i < ngrp; i += BLOCK) lgrp[i] = 0;
    for (int i = threadIdx.x; i < ntab; i += BLOCK) ltab[i] = D->tab[i];
    const int np = D->npred;
    for (int i = threadIdx.x; i < np; i += BLOCK) {
      const BatchPred P = D->pred[i];
      H.pred[i] = make_int4(P.lo, (int)((unsigned)P.hi - (unsigned)P.lo), P.q, P.col);
    }
    if (threadIdx.x < kBatchMaxCols) H.col[threadIdx.x] = D->col[threadIdx.x];
    if (threadIdx.x <= kBatchMaxCols) H.pred_begin[threadIdx.x] = D->pred_begin[threadIdx.x];
    if (threadIdx.x < kBatchMaxExpr) {
      const BatchExpr E = D->expr[threadIdx.x];
      H.expr[threadIdx.x] = make_int4(E.kind, E.a1, E.a2, 0);
    }
    if (threadIdx.x < kBatchMaxIvl) H.ivl[threadIdx.x] = D->ivl[threadIdx.x];
    if (threadIdx.x < 4) {
      const int d = threadIdx.x;
      const BatchDim& B = D->dim[d];
      BatchHotDim& Hd = H.dim[d];
      Hd.mk = B.mk;
      Hd.pk = B.pk;
      Hd.uses = B.uses;
      Hd.groupers = B.groupers;
      Hd.len = B.len;
      Hd.minkey = B.minkey;
      Hd.pw = B.pw;
      Hd.form = B.form;
      Hd.nattr = B.nattr;
      Hd.mstride = B.mstride;
      Hd.cb = B.cb;
      Hd.dict = B.dict_off;
      for (int a = 0; a < kBatchMaxAttr; a++) {
        // the attribute and dictionary forms store value - base + 1
        Hd.base[a] = B.attr[a].base - (B.form != kDimMask ? 1 : 0);
        Hd.off[a] = B.attr[a].off;
        Hd.bits[a] = B.attr[a].bits;
        Hd.lut[a] = B.attr[a].lut;
      }
    }
    if (threadIdx.x == 0) {
      int cm = 0, um = 0;
      for (int c = 0; c < D->ncols; c++)
        if (D->col[c]) cm |= 1 << c;
      for (int d = 0; d < 4; d++)
        if (D->dim[d].used) um |= 1 << d;
      H.ncols = D->ncols;
      H.colmask = cm;
      H.usedmask = um;
      H.nexpr = D->nexpr;
      H.npred = np;
      H.nq = nq;
      H.ntab = ntab;
      H.nivl = D->nivl;
      H.attr16 = D->attr16;
    }
  }
  __syncthreads();

  const int colmask = __builtin_amdgcn_readfirstlane(H.colmask);
  const int usedmask = __builtin_amdgcn_readfirstlane(H.usedmask);
  const int nexpr = __builtin_amdgcn_readfirstlane(H.nexpr);
  const int nivl = __builtin_amdgcn_readfirstlane(H.nivl);
  const bool attr16 = __builtin_amdgcn_readfirstlane(H.attr16) != 0;
  const unsigned long long all = D->all;
  const int lane = threadIdx.x & (kWave - 1);

  const int num_tuples = D->num_tuples;
  const short* __restrict__ seg_group = L.seg_group + D->seg_begin;
  const int ntiles = (num_tuples + TILE - 1) / TILE;
  const int per = (ntiles + gridDim.x - 1) / gridDim.x;
  const int t_begin = blockIdx.x * per;
  const int t_end = min(ntiles, t_begin + per);

  int cur_seg = -1;
  int slotv = 0;  // lane c: cache slot of column c in segment cur_seg
  auto load_cols = [&](int t, int (&x)[MAXC][IPT]) {
    const int pos = t * TILE;
    const int sidx = pos >> kSegLog;
    if (sidx != cur_seg) {
      const int seg = __builtin_amdgcn_readfirstlane((int)seg_group[sidx]);
      const int* tab = lane < MAXC && (colmask >> lane & 1) ? H.col[lane] : nullptr;
      slotv = tab ? gload<int>(tab, seg) : 0;
      cur_seg = sidx;
    }
    const int off = pos & kSegMask;
    const int n = min(TILE, num_tuples - pos);
    // every slot into a scalar first, then the loads back to back; a slot the
    // batch does not read points at the first column it does read (its tile is
    // loaded anyway; the values are never used), so that no load sits under a
    // per-column branch, which would make each load wait for the previous one
    const int c0 = __builtin_ctz((unsigned)colmask | (1u << (MAXC - 1)));
    const int slot0 = __builtin_amdgcn_readlane(slotv, c0);
    const int* base[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; c++) {
      const int sl = (colmask >> c & 1) ? __builtin_amdgcn_readlane(slotv, c) : slot0;
      base[c] = L.cache + ((size_t)(unsigned)sl << kSegLog) + off;
    }
    if (n == TILE) {
#pragma unroll
      for (int c = 0; c < MAXC; c++) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
          const v4i t = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(base[c]) + v * BLOCK + threadIdx.x);
          x[c][4 * v + 0] = t.x; x[c][4 * v + 1] = t.y; x[c][4 * v + 2] = t.z; x[c][4 * v + 3] = t.w;
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < MAXC; c++) {
        if (colmask >> c & 1) {
          load_tile<BLOCK, VPT>(base[c], false, n, x[c]);
        } else {
#pragma unroll
          for (int j = 0; j < IPT; j++) x[c][j] = 0;
        }
      }
    }
  };

  unsigned long long alive[IPT];
  uint32_t attrw[4][IPT];
  int ev[kBatchMaxExpr][IPT];

  // the alive mask of every item, its group attribute words and the values
  // of the aggregate expressions
  auto masks = [&](int t, const int (&x)[MAXC][IPT]) {
    const int pos = t * TILE;
    const int n = min(TILE, num_tuples - pos);
    const uint32_t inr = n == TILE ? (IPT == 32 ? 0xFFFFFFFFu : (1u << IPT) - 1u) : range_mask<BLOCK, VPT>(n);
#pragma unroll
    for (int j = 0; j < IPT; j++) alive[j] = (inr >> j & 1u) ? all : 0ull;

    // fact predicates with an interval table: bucket -> rank at the bucket
    // start -> at most kmax breakpoints inside it -> the interval's mask
    for (int v = 0; v < nivl; v++) {
      const BatchIvl& I = H.ivl[v];
      const int col = __builtin_amdgcn_readfirstlane(I.col);
      const int base = __builtin_amdgcn_readfirstlane(I.base);
      const int shift = __builtin_amdgcn_readfirstlane(I.shift);
      const unsigned last = (unsigned)__builtin_amdgcn_readfirstlane(I.nb) - 1u;
      const int kmax = __builtin_amdgcn_readfirstlane(I.kmax);
      const int nbp = __builtin_amdgcn_readfirstlane(I.n);
      const unsigned short* rk = reinterpret_cast<const unsigned short*>(ltab + __builtin_amdgcn_readfirstlane(I.rank_off));
      const int* bp = reinterpret_cast<const int*>(ltab + __builtin_amdgcn_readfirstlane(I.bp_off));
      const unsigned long long* mt =
          reinterpret_cast<const unsigned long long*>(ltab + __builtin_amdgcn_readfirstlane(I.mask_off));
      int xv[IPT];
#pragma unroll
      for (int j = 0; j < IPT; j++) xv[j] = 0;
#pragma unroll
      for (int c = 0; c < MAXC; c++)
        if (c == col) {
#pragma unroll
          for (int j = 0; j < IPT; j++) xv[j] = x[c][j];
        }
      int r[IPT];
#pragma unroll
      for (int j = 0; j < IPT; j++) r[j] = rk[min((unsigned)(xv[j] - base) >> shift, last)];
      int add[IPT];
#pragma unroll
      for (int j = 0; j < IPT; j++) add[j] = 0;
      for (int k = 0; k < kmax; k++) {
#pragma unroll
        for (int j = 0; j < IPT; j++) add[j] += xv[j] >= bp[r[j] + k];
      }
#pragma unroll
      for (int j = 0; j < IPT; j++) alive[j] &= mt[min(r[j] + add[j], nbp)];
    }
    // the other fact BETWEEN predicates (GPUProcessing.h:72-86), grouped by
    // column; the half a predicate clears is chosen by a wave-uniform branch
#pragma unroll
    for (int c = 0; c < MAXC; c++) {
      if (!(colmask >> c & 1)) continue;
      const int pb = __builtin_amdgcn_readfirstlane(H.pred_begin[c]);
      const int pe = __builtin_amdgcn_readfirstlane(H.pred_begin[c + 1]);
      for (int p = pb; p < pe; p++) {
        const int4 P = H.pred[p];
        const int lo = __builtin_amdgcn_readfirstlane(P.x);
        const unsigned span = (unsigned)__builtin_amdgcn_readfirstlane(P.y);
        const int q = __builtin_amdgcn_readfirstlane(P.z);
        const unsigned long long keep = ~(1ull << q);
#pragma unroll
        for (int j = 0; j < IPT; j++)
          if ((unsigned)(x[c][j] - lo) > span) alive[j] &= keep;
      }
    }

    // one probe per dimension for all queries: mask form, the slot's query
    // mask (and, in a 16-byte entry, its group attribute word); attribute
    // form, the slot's attribute word, whose filtered attributes index their
    // mask tables in LDS.  The gathers of dimension d+1 are issued before
    // dimension d is used, and the loads of one dimension sit under no
    // per-item branch: a load consumed before the next is issued waits for
    // the loads before it.  An item no query needs, or whose key has no
    // slot, reads slot len, which holds zero (no query passes); an unused
    // dimension reads 8 bytes of the descriptor.  An attribute entry is pw
    // bytes (1, 2 or 4) at slot x pw, so it never straddles the aligned
    // 8-byte word holding it.
    const void* const dummy = D;
    unsigned long long mv[2][IPT];  // two dimensions in flight
    uint32_t aw[2][IPT];            // (16-byte mask entries: their attribute word)
    // the entry's bit offset in the loaded 8 bytes (dictionary and attribute
    // forms), one byte per item
    uint32_t shp[2][(IPT + 3) / 4];
    // the slot an item reads: its own, or slot len (zero) when no query
    // needs it or its key has no slot
    auto slot_of = [&](int d, int j, int len, int mk, unsigned long long uses) {
      const int h = phash(x[d][j], len, mk);
      return (unsigned)h < (unsigned)len && (alive[j] & uses) ? (unsigned)h : (unsigned)len;
    };
    // byte offset of the load holding the entry of slot s, and the entry's
    // bit offset in the loaded 8 bytes: mask form 8 or 16 bytes per slot,
    // attribute form pw bytes (in the aligned 8-byte word), dictionary form
    // cb bits (the 8 bytes from the aligned 4-byte word holding its first bit;
    // cb <= 11 bits, so they hold it)
    auto entry_at = [](int form, unsigned stride, unsigned s, unsigned& sh) -> uint32_t {
      if (form == kDimDict) {
        const uint32_t bit = s * stride;
        const uint32_t at = (bit >> 3) & ~3u;
        sh = bit - 8u * at;
        return at;
      }
      const uint32_t b = s * stride;  // (a table is far below 4 GB)
      sh = (b & 7u) * 8u;
      return b & ~7u;
    };
    auto issue = [&](int d) {
      const BatchHotDim& Hd = H.dim[d];
      const bool used = usedmask >> d & 1;
      const int len = used ? __builtin_amdgcn_readfirstlane(Hd.len) : 0;
      const int mk = __builtin_amdgcn_readfirstlane(Hd.minkey);
      const int form = __builtin_amdgcn_readfirstlane(Hd.form);
      const bool wide = used && form == kDimMask && __builtin_amdgcn_readfirstlane(Hd.mstride) == 2;
      const unsigned stride = wide ? 16u
                              : form == kDimMask ? 8u
                              : (unsigned)__builtin_amdgcn_readfirstlane(form == kDimDict ? Hd.cb : Hd.pw);
      const unsigned long long uses = Hd.uses;
      const void* tabp = uniform_ptr(!used ? dummy : form == kDimMask ? static_cast<const void*>(Hd.mk) : Hd.pk);
      uint32_t b[IPT];
#pragma unroll
      for (int w = 0; w < (IPT + 3) / 4; w++) shp[d & 1][w] = 0u;
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        unsigned sh = 0;
        b[j] = wide ? slot_of(d, j, len, mk, uses) * 16u : entry_at(form, stride, slot_of(d, j, len, mk, uses), sh);
        shp[d & 1][j >> 2] |= sh << (8 * (j & 3));
      }
      if (wide) {
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          const v4i t = gload_at<v4i>(tabp, b[j]);
          mv[d & 1][j] = (unsigned long long)(unsigned)t.y << 32 | (unsigned)t.x;
          aw[d & 1][j] = (uint32_t)t.z;
        }
      } else {
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          mv[d & 1][j] = gload_at<unsigned long long>(tabp, b[j]);
          aw[d & 1][j] = 0u;
        }
      }
    };
    auto consume = [&](int d) {
#pragma unroll
      for (int j = 0; j < IPT; j++) attrw[d][j] = 0;
      if (!(usedmask >> d & 1)) return;
      const BatchHotDim& Hd = H.dim[d];
      const unsigned long long uses = Hd.uses;
      const unsigned long long(&m)[IPT] = mv[d & 1];
      const int form = __builtin_amdgcn_readfirstlane(Hd.form);
      if (form == kDimMask) {
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          alive[j] &= m[j] | ~uses;
          attrw[d][j] = aw[d & 1][j];
        }
        return;
      }
      // the entry's bit offset as issued (an item that has lost every query
      // using the dimension since then reads bits no query uses)
      auto sh_of = [&](int j) { return (shp[d & 1][j >> 2] >> (8 * (j & 3))) & 0xFFu; };
      if (form == kDimDict) {
        const unsigned cb = (unsigned)__builtin_amdgcn_readfirstlane(Hd.cb);
        const uint4* dict = reinterpret_cast<const uint4*>(ltab + __builtin_amdgcn_readfirstlane(Hd.dict));
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          const unsigned sh = sh_of(j);
          const uint32_t code = (uint32_t)(m[j] >> sh) & ((1u << cb) - 1u);
          const uint4 t = dict[code];
          alive[j] &= ((unsigned long long)t.y << 32 | t.x) | ~uses;
          attrw[d][j] = t.z;
        }
        return;
      }
      const unsigned pw = (unsigned)__builtin_amdgcn_readfirstlane(Hd.pw);
      const uint32_t wm = pw >= 4 ? 0xFFFFFFFFu : (1u << (8 * pw)) - 1u;
#pragma unroll
      for (int j = 0; j < IPT; j++) attrw[d][j] = (uint32_t)(m[j] >> sh_of(j)) & wm;
      const int na = __builtin_amdgcn_readfirstlane(Hd.nattr);
      for (int a = 0; a < na; a++) {
        const int lut = __builtin_amdgcn_readfirstlane(Hd.lut[a]);
        if (lut < 0) continue;
        const unsigned off = (unsigned)__builtin_amdgcn_readfirstlane(Hd.off[a]);
        const unsigned bits = (unsigned)__builtin_amdgcn_readfirstlane(Hd.bits[a]);
        const unsigned long long* lt = reinterpret_cast<const unsigned long long*>(ltab + lut);
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          const uint32_t f = __builtin_amdgcn_ubfe(attrw[d][j], off, bits);
          alive[j] &= lt[f ? f - 1u : 0u];  // (an empty slot clears its users below)
        }
      }
#pragma unroll
      for (int j = 0; j < IPT; j++)
        if (!attrw[d][j]) alive[j] &= ~uses;
    };
    // (the probes see the alive masks from before the previous dimension)
    issue(0);
#pragma unroll
    for (int d = 0; d < 4; d++) {
      if (d + 1 < 4) issue(d + 1);
      consume(d);
    }
    // 16-bit attribute words: dimensions 1 and 3 into the high halves of 0
    // and 2 (one ds_bpermute per two dimensions in the pair loop)
    if (attr16) {
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        attrw[0][j] |= attrw[1][j] << 16;
        attrw[2][j] |= attrw[3][j] << 16;
      }
    }
    // aggregate expressions of the batch (int32 arithmetic, GPUProcessing.h:149, 1763)
#pragma unroll
    for (int e = 0; e < kBatchMaxExpr; e++) {
#pragma unroll
      for (int j = 0; j < IPT; j++) ev[e][j] = 0;
      if (e >= nexpr) continue;
      const int4 E = H.expr[e];
      const int kind = __builtin_amdgcn_readfirstlane(E.x);
      const int a1 = __builtin_amdgcn_readfirstlane(E.y);
      const int a2 = __builtin_amdgcn_readfirstlane(E.z);
      int v1[IPT], v2[IPT];
#pragma unroll
      for (int j = 0; j < IPT; j++) v1[j] = v2[j] = 0;
#pragma unroll
      for (int c = 0; c < MAXC; c++) {
        if (c == a1) {
#pragma unroll
          for (int j = 0; j < IPT; j++) v1[j] = x[c][j];
        }
        if (c == a2) {
#pragma unroll
          for (int j = 0; j < IPT; j++) v2[j] = x[c][j];
        }
      }
#pragma unroll
      for (int j = 0; j < IPT; j++)
        ev[e][j] = kind == kExprProduct ? (int)((unsigned)v1[j] * (unsigned)v2[j])
                                        : (int)((unsigned)v1[j] - (unsigned)v2[j]);
    }
  };

  // Queries that store their keys in the pair loop (their add must return
  // the previous sum, with the row's group values at hand) keep a per-lane
  // loop of their own; every other pair goes through the wave's pair list.
  unsigned long long keysq = 0;
  for (int q = 0; q < nq; q++)
    if (query(q, true).flags & kQueryKeysInLoop) keysq |= 1ull << q;

  auto group_of = [&](const BatchQuery& Q, const uint32_t (&a)[4]) {
    // field d's shift and width are bytes d of two words; v_bfe_u32 reads only
    // the low 5 bits of its offset and width operands, so byte d needs no mask
    uint32_t S, B;
    __builtin_memcpy(&S, Q.shift, 4);
    __builtin_memcpy(&B, Q.bits, 4);
    return mod_pos(Q.hbase + (int)__umul24(__builtin_amdgcn_ubfe(a[0], S, B), (unsigned)Q.uniq[0]) +
                       (int)__umul24(__builtin_amdgcn_ubfe(a[1], S >> 8, B >> 8), (unsigned)Q.uniq[1]) +
                       (int)__umul24(__builtin_amdgcn_ubfe(a[2], S >> 16, B >> 16), (unsigned)Q.uniq[2]) +
                       (int)__umul24(__builtin_amdgcn_ubfe(a[3], S >> 24, B >> 24), (unsigned)Q.uniq[3]),
                   Q.total_val);
  };
  // (four named values, not an array: a select chain over an array indexed
  // by the query's expression becomes a dynamic index into scratch memory)
  static_assert(kBatchMaxExpr == 4, "value_of selects among four expressions");
  auto value_of = [&](const BatchQuery& Q, int v0, int v1, int v2, int v3) {
    const unsigned x = Q.expr;
    int val = v0;
    val = x == 1u ? v1 : val;
    val = x == 2u ? v2 : val;
    val = x == 3u ? v3 : val;
    return val;
  };

  // every surviving (row, query) pair.  An index outside [0, total_val)
  // (group values below min_val) would write outside the query's result; the
  // reference does, here the pair is dropped.  (Scoped atomics keep the LDS
  // ones ds_ instructions: one atomicAdd on either pointer would be merged
  // into a slower flat atomic.)
  //
  // Per row slot j the lanes list their pairs in the wave's pair list (each
  // lane's entries at an offset from one returning add on the wave's
  // counter), then every lane aggregates one listed pair per step: a wave
  // takes ceil(pairs / 64) steps instead of the largest per-lane pair count.
  // The row's attribute words and expression values come from the pair's
  // source lane by ds_bpermute, which every lane of the wave executes (a
  // disabled source lane would read as zero).  No add returns a value except
  // the 32-bit LDS sums' (the carry): keys are written after the scan from
  // the touched bitmaps.
  // (the wave's pair list: a counter word, then kBatchPairCap entries, at the
  // end of the dynamic LDS; the counter is zero between rounds)
  unsigned* const lpn = reinterpret_cast<unsigned*>(lds_raw + D->lds_dyn - batch_pair_bytes(BLOCK)) +
                        (threadIdx.x >> 6) * (1 + kBatchPairCap / 2);
  unsigned short* const lpl = reinterpret_cast<unsigned short*>(lpn + 1);
  if (lane == 0) *lpn = 0u;
  // (A16: the batch's 16-bit attribute words are packed two per register, a
  // compile-time choice so that no copy of a ds_bpermute result sits between
  // the step's LDS reads)
  auto pairs = [&](auto A16, auto U16) {
    constexpr bool kA16 = decltype(A16)::value;
    constexpr bool kU16 = decltype(U16)::value;
    // the device address of query q's result words
    auto res_of = [&](int q) {
      const uint4 r = lqc[(kU16 ? 2 * kBatchMaxQ : 0) + q];
      return reinterpret_cast<int*>((unsigned long long)r.y << 32 | r.x);
    };
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      unsigned long long m = alive[j] & ~keysq;
      while (__ballot(m != 0ull)) {
        // list: this lane's pairs at [p, p + popcount); entries past the
        // list's end stay in m for the next round
        unsigned p = 0;
        if (m) p = __hip_atomic_fetch_add(lpn, (unsigned)__popcll(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        for (; m && p < (unsigned)kBatchPairCap; p++) {
          const int q = __builtin_ctzll(m);
          m &= m - 1;
          lpl[p] = (unsigned short)(q << 8 | lane << 2);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int n = min(__builtin_amdgcn_readfirstlane(
                              (int)__hip_atomic_load(lpn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT)),
                          kBatchPairCap);
        for (int b = 0; b < n; b += kWave) {
          // (entries at and past n are stale or unwritten: read, never used)
          const unsigned ent = lpl[b + lane];
          const int src = (int)(ent & 0xFCu);
          uint32_t a[4];
          a[0] = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)attrw[0][j]);
          a[2] = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)attrw[2][j]);
          if constexpr (kA16) {
            a[1] = a[0];
            a[3] = a[2];
          } else {
            a[1] = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)attrw[1][j]);
            a[3] = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)attrw[3][j]);
          }
          const int v0 = __builtin_amdgcn_ds_bpermute(src, ev[0][j]);
          const int v1 = nexpr > 1 ? __builtin_amdgcn_ds_bpermute(src, ev[1][j]) : 0;
          const int v2 = nexpr > 2 ? __builtin_amdgcn_ds_bpermute(src, ev[2][j]) : 0;
          const int v3 = nexpr > 3 ? __builtin_amdgcn_ds_bpermute(src, ev[3][j]) : 0;
          if (b + lane >= n) continue;
          // the descriptor's operands: two 16-byte LDS reads with uniq16
          // (uniq as 16-bit halves), else three (c0 = res, total_val, acc;
          // c1 = hbase, expr | sink << 8 | touch << 16, shift bytes, bits
          // bytes; c2 = uniq)
          const int q = (int)(ent >> 8);
          unsigned total, accb, hb, ess, sh, bt, u0, u1, u2, u3;
          if constexpr (kU16) {
            const uint4 c0 = lqc[q], c1 = lqc[kBatchMaxQ + q];
            total = c0.x, accb = c0.y, hb = c0.z, ess = c0.w, sh = c1.x, bt = c1.y;
            u0 = c1.z & 0xFFFFu, u1 = c1.z >> 16, u2 = c1.w & 0xFFFFu, u3 = c1.w >> 16;
          } else {
            const uint4 c0 = lqc[q], c1 = lqc[kBatchMaxQ + q], c2 = lqc[2 * kBatchMaxQ + q];
            total = c0.z, accb = c0.w, hb = c1.x, ess = c1.y, sh = c1.z, bt = c1.w;
            u0 = c2.x, u1 = c2.y, u2 = c2.z, u3 = c2.w;
          }
          const int h = mod_pos((int)hb + (int)__umul24(__builtin_amdgcn_ubfe(a[0], sh, bt), u0) +
                                    (int)__umul24(__builtin_amdgcn_ubfe(a[1], sh >> 8, bt >> 8), u1) +
                                    (int)__umul24(__builtin_amdgcn_ubfe(a[2], sh >> 16, bt >> 16), u2) +
                                    (int)__umul24(__builtin_amdgcn_ubfe(a[3], sh >> 24, bt >> 24), u3),
                                (int)total);
          if ((unsigned)h >= total) continue;
          const unsigned expr = ess & 0xFFu, sink = (ess >> 8) & 0xFFu, touch = ess >> 16;
          int val = v0;
          val = expr == 1u ? v1 : val;
          val = expr == 2u ? v2 : val;
          val = expr == 3u ? v3 : val;
          unsigned char* const acc = lds_raw + accb;
          auto gsum_of = [&](int qq) {
            return (__attribute__((address_space(1))) unsigned long long*)(res_of(qq) + 6 * (size_t)h + 4);
          };
          if (sink == kSinkLds64) {
            (void)__hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(acc) + h,
                                         (unsigned long long)(long long)val, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
          } else if (sink == kSinkLds32) {
            // 32-bit LDS sums: (int64)val = u - 2^32 [val < 0]; the low word
            // wraps when old + u carries, and the 2^32 that leaves it goes to
            // the result
            const unsigned u = (unsigned)val;
            const unsigned old = __hip_atomic_fetch_add(reinterpret_cast<unsigned*>(acc) + h, u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            const int delta = (int)(old + u < old) - (int)(val < 0);
            if (delta) {
              const unsigned long long carry = (unsigned long long)(long long)delta << 32;
              (void)__hip_atomic_fetch_add(gsum_of(q), carry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              // identical queries sharing these sums take the carry too (their
              // low words come from the same LDS sums at the flush)
              for (int al = (int)lqc[3 * kBatchMaxQ + q].w; al >= 0; al = (int)lqc[3 * kBatchMaxQ + al].w)
                (void)__hip_atomic_fetch_add(gsum_of(al), carry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          } else {
            (void)__hip_atomic_fetch_add(gsum_of(q), (unsigned long long)(long long)val, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
          }
          if (touch != kBatchNoTouch)
            (void)__hip_atomic_fetch_or(reinterpret_cast<unsigned*>(lds_raw) + touch + (h >> 5), 1u << (h & 31),
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // (every lane has read the list and the counter: LDS operations of a
        // wave complete in order)
        __hip_atomic_store(lpn, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
      uint32_t a[4];
      a[0] = attrw[0][j];
      a[1] = kA16 ? attrw[0][j] : attrw[1][j];
      a[2] = attrw[2][j];
      a[3] = kA16 ? attrw[2][j] : attrw[3][j];
      m = alive[j] & keysq;
      while (m) {
        const int q = __builtin_ctzll(m);
        m &= m - 1;
        const BatchQuery Q = query(q, true);
        const int h = group_of(Q, a);
        if ((unsigned)h >= (unsigned)Q.total_val) continue;
        const int val = value_of(Q, ev[0][j], ev[1][j], ev[2][j], ev[3][j]);
        const unsigned long long add = (unsigned long long)(long long)val;
        int* res = reinterpret_cast<int*>(Q.res) + 6 * (size_t)h;
        auto* gsum = (__attribute__((address_space(1))) unsigned long long*)reinterpret_cast<unsigned long long*>(res + 4);
        unsigned char* const acc = lds_raw + Q.acc;
        // keys that do not decode from the index: stored whenever the sum the
        // add returns was zero -- at the first pair of the group in this
        // workgroup (LDS) or overall (global), and again only if a running sum
        // passes through zero.  The reference stores them for every row of the
        // group (GPUProcessing.h:144-147); they are equal, so any store suffices.
        unsigned long long old;
        if (Q.sink == kSinkLds64)
          old = __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(acc) + h, add, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        else
          old = __hip_atomic_fetch_add(gsum, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(Q.flags & kQueryAggr) && old == 0ull) {
          int gv[4];
#pragma unroll
          for (int d = 0; d < 4; d++)
            gv[d] = Q.attr[d] < 0 ? 0
                                  : H.dim[d].base[Q.attr[d]] + (int)__builtin_amdgcn_ubfe(a[d], Q.shift[d], Q.bits[d]);
          int2* key = reinterpret_cast<int2*>(res);
          key[0] = make_int2(gv[0], gv[1]);
          key[1] = make_int2(gv[2], gv[3]);
        }
      }
    }
  };

  if (t_begin < t_end) {
    int x[MAXC][IPT];
    load_cols(t_begin, x);
    for (int t = t_begin; t < t_end; t++) {
      masks(t, x);
      if (t + 1 < t_end) load_cols(t + 1, x);
      if (attr16) {
        if (lq16) pairs(std::true_type{}, std::true_type{});
        else pairs(std::true_type{}, std::false_type{});
      } else {
        if (lq16) pairs(std::false_type{}, std::true_type{});
        else pairs(std::false_type{}, std::false_type{});
      }
    }
  }

  // flush: the touched bitmaps into the query's global bitmap (read by the
  // keys kernel), the LDS sums of the touched groups into the result (adding
  // a zero sum changes nothing)
  __syncthreads();
  unsigned* const touch_arena = D->touch_arena;
  for (int q = 0; q < nq; q++) {
    if (!(all >> q & 1ull)) continue;
    const BatchQuery Q = query(q, true);
    const int G = Q.total_val, flags = Q.flags, sink = Q.sink;
    const int tw = batch_touch_words(G, flags);  // (Q.touch is kBatchNoTouch iff tw == 0)
    const unsigned* tb = reinterpret_cast<const unsigned*>(lds_raw) + (tw ? Q.touch : 0);
    const unsigned* acc = reinterpret_cast<const unsigned*>(lds_raw + Q.acc);
    auto* res = reinterpret_cast<unsigned long long*>(Q.res);
    // (identical queries chained from Q.alias take the same sums and touched
    // groups: Q's LDS sums count this launch's rows, which are theirs too)
    auto add = [&](int h) {
      const unsigned long long s = sink == kSinkLds64 ? reinterpret_cast<const unsigned long long*>(acc)[h]
                                                      : (unsigned long long)acc[h];
      if (!s) return;
      atomicAdd(res + 3 * (size_t)h + 2, s);
      for (int a = Q.alias; a >= 0;) {
        const BatchQuery A = query(a, true);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.res) + 3 * (size_t)h + 2, s);
        a = A.alias;
      }
    };
    if (tw == 0) {
      if (sink != kSinkGlobal)
        for (int h = threadIdx.x; h < G; h += BLOCK) add(h);
      continue;
    }
    for (int w = threadIdx.x; w < tw; w += BLOCK) {
      unsigned bits = tb[w];
      if (!bits) continue;
      (void)__hip_atomic_fetch_or(touch_arena + Q.gtouch + w, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int a = Q.alias; a >= 0;) {
        const BatchQuery A = query(a, true);
        (void)__hip_atomic_fetch_or(touch_arena + A.gtouch + w, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a = A.alias;
      }
      if (sink == kSinkGlobal) continue;
      while (bits) {
        const int b = __builtin_ctz(bits);
        bits &= bits - 1u;
        add(32 * w + b);
      }
    }
  }
}

namespace {

template <int BLOCK, int VPT, int MAXC>
const void* scan_fn() {
  return reinterpret_cast<const void*>(&batch_kernel<BLOCK, VPT, MAXC>);
}

// Scan kernel of a configuration: columns held in registers per tile = 8 when
// the batch reads at most 8 fact columns, else 12.  With 12 columns the
// kernel keeps 4 rows per thread and at most 512 threads per workgroup (the
// register budget of 1,024 threads, or of 8 rows per thread, would spill);
// the configuration's LDS budget is kept.
const void* pick(const BatchConfig& c, int ncols) {
  const bool small = ncols <= 8;
  if (c.vpt != 1 && c.vpt != 2) return nullptr;
  if (c.block != 256 && c.block != 512 && (c.block != 1024 || c.vpt != 1)) return nullptr;
  if (!small) return c.block == 256 ? scan_fn<256, 1, 12>() : scan_fn<512, 1, 12>();
  if (c.vpt == 1) {
    if (c.block == 256) return scan_fn<256, 1, 8>();
    if (c.block == 512) return scan_fn<512, 1, 8>();
    return scan_fn<1024, 1, 8>();
  }
  return c.block == 256 ? scan_fn<256, 2, 8>() : scan_fn<512, 2, 8>();
}

template <int BLOCK, int VPT, int MAXC>
hipError_t launch(const BatchLaunch& L, int lds, hipStream_t s) {
  constexpr int TILE = BLOCK * 4 * VPT;
  const auto f = &batch_kernel<BLOCK, VPT, MAXC>;
  const int grid = balanced_grid((L.max_tuples + TILE - 1) / TILE,
                                 resident_groups(reinterpret_cast<const void*>(f), BLOCK, lds - sizeof(BatchHot)));
  f<<<grid, BLOCK, lds - (int)sizeof(BatchHot), s>>>(L);
  return hipGetLastError();
}

hipError_t launch_c8(const BatchLaunch& L, const BatchConfig& c, hipStream_t s) {
  if (c.vpt == 1) {
    if (c.block == 256) return launch<256, 1, 8>(L, c.lds_bytes, s);
    if (c.block == 512) return launch<512, 1, 8>(L, c.lds_bytes, s);
    return launch<1024, 1, 8>(L, c.lds_bytes, s);
  }
  if (c.block == 256) return launch<256, 2, 8>(L, c.lds_bytes, s);
  return launch<512, 2, 8>(L, c.lds_bytes, s);
}

hipError_t launch_c12(const BatchLaunch& L, const BatchConfig& c, hipStream_t s) {
  return c.block == 256 ? launch<256, 1, 12>(L, c.lds_bytes, s) : launch<512, 1, 12>(L, c.lds_bytes, s);
}

}  // namespace

bool batch_config_ok(const BatchConfig& c) {
  return pick(c, 8) != nullptr && c.lds_bytes >= (int)sizeof(BatchHot) + batch_pair_bytes(c.block) &&
         c.lds_bytes <= 160 * 1024;
}

hipError_t batch_prepare(const BatchConfig& c, int ncols) {
  const void* f = pick(c, ncols);
  if (!f) return hipErrorInvalidValue;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                             std::max(c.lds_bytes - (int)sizeof(BatchHot), 1));
}

int batch_resident(const BatchConfig& c, int ncols) {
  const void* f = pick(c, ncols);
  const int block = ncols <= 8 ? c.block : std::min(c.block, 512);
  return f ? resident_groups(f, block, c.lds_bytes - sizeof(BatchHot)) : 0;
}

hipError_t batch_fill(const BatchLaunch& L, hipStream_t s) {
  batch_fill_kernel<<<2048, 256, 0, s>>>(L.desc, L.status);
  return hipGetLastError();
}

hipError_t batch_records(const BatchLaunch& L, int d, int nrows, hipStream_t s) {
  const int grid = std::max(1, std::min(4096, (nrows + 255) / 256));
  batch_records_kernel<<<grid, 256, 0, s>>>(L, d);
  return hipGetLastError();
}

hipError_t batch_keys(const BatchLaunch& L, hipStream_t s) {
  batch_keys_kernel<<<dim3(32, kBatchMaxQ), 256, 0, s>>>(L.desc);
  return hipGetLastError();
}

hipError_t batch_scan(const BatchLaunch& L, const BatchConfig& c, int ncols, hipStream_t s) {
  if (!batch_config_ok(c) || ncols > kBatchMaxCols) return hipErrorInvalidValue;
  return ncols <= 8 ? launch_c8(L, c, s) : launch_c12(L, c, s);
}

}  // namespace mordred
