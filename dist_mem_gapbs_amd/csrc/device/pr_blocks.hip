// Host-side construction of the propagation layout of a PE block's
// compensated PageRank iteration (the block's rows in the caller's order,
// sources anywhere in [0, n_gather)).
//
// Sources are cut into windows of kBWin vertices (one window of
// contributions fills a workgroup's LDS); rows into buckets of at most
// kBRows consecutive rows and kBEdges in-edges (a row with more in-edges is
// a bucket of its own, summed in units of at most kBEdges edges).  Edge
// (u -> r) belongs to segment (window u / kBWin, bucket of r).
//
//   window-major array (pass 1 reads it): per window its edges' 16-bit
//     source offsets, segment by segment in bucket order, the window's
//     start padded to a multiple of kBPer edges;
//   bucket-major array (pass 1 writes it, pass 2 reads it): per bucket its
//     edges' contributions, segment by segment in window order, and beside
//     it (static) each edge's row offset inside its bucket.
// The i-th edge of a segment sits at the i-th place of the segment in both
// arrays.  Per window, the segment table gives each non-empty segment's
// start inside the window and its destination in the bucket-major array
// (stored as destination - start, so an edge's destination is that plus
// its index in the window).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../host/parallel.h"
#include "pr.h"

namespace dmm {

int build_blocks_host(int64_t n, int64_t n_gather, const int64_t* off, const int32_t* nbr, BlockHost* out,
                      std::string* err) {
  const int nt = resolve_threads(0);
  const int64_t m = off[n];
  if (m >= (int64_t(1) << 31) || n_gather >= (int64_t(1) << 31)) {
    *err = "PE block too large for the propagation layout";
    return DMM_ERR_CAPACITY;
  }
  const int32_t nwin = static_cast<int32_t>(std::max<int64_t>(1, (n_gather + kBWin - 1) / kBWin));
  // 1. Buckets.
  std::vector<int32_t> brow{0};
  {
    int64_t rows = 0, edges = 0;
    for (int64_t r = 0; r < n; r++) {
      const int64_t d = off[r + 1] - off[r];
      if (rows > 0 && (rows == kBRows || edges + d > kBEdges)) {
        brow.push_back(static_cast<int32_t>(r));
        rows = edges = 0;
      }
      rows++;
      edges += d;
    }
    if (n > 0) brow.push_back(static_cast<int32_t>(n));
  }
  const int64_t nb = static_cast<int64_t>(brow.size()) - 1;
  if (nb > kBSegMax) {
    *err = "PE block has more row buckets than a window's segment table holds";
    return DMM_ERR_CAPACITY;
  }
  // 2. Edges per (window, bucket), cnt[w * nb + b].
  std::vector<int32_t> cnt(static_cast<size_t>(nwin) * std::max<int64_t>(1, nb), 0);
  parallel_dynamic(nb, nt, 16, [&](int64_t b) {
    for (int64_t e = off[brow[b]]; e < off[brow[b + 1]]; e++) {
      const int32_t u = nbr[e];
      cnt[static_cast<size_t>(u / kBWin) * nb + b]++;
    }
  });
  // 3. Window-major starts (padded) and bucket-major starts per segment.
  std::vector<int64_t> wstart(nwin + 1, 0);
  std::vector<int32_t> wlen(nwin, 0);
  for (int32_t w = 0; w < nwin; w++) {
    int64_t len = 0;
    for (int64_t b = 0; b < nb; b++) len += cnt[static_cast<size_t>(w) * nb + b];
    wlen[w] = static_cast<int32_t>(len);
    wstart[w + 1] = wstart[w] + (len + kBPer - 1) / kBPer * kBPer;
  }
  std::vector<int64_t> bstart(nb + 1, 0);
  for (int64_t b = 0; b < nb; b++) bstart[b + 1] = bstart[b] + (off[brow[b + 1]] - off[brow[b]]);
  // Segment (w, b): start in the window-major array, destination in the
  // bucket-major array.
  std::vector<int64_t> wpos(static_cast<size_t>(nwin) * std::max<int64_t>(1, nb)),
      bpos(static_cast<size_t>(nwin) * std::max<int64_t>(1, nb));
  for (int32_t w = 0; w < nwin; w++) {
    int64_t p = wstart[w];
    for (int64_t b = 0; b < nb; b++) {
      wpos[static_cast<size_t>(w) * nb + b] = p;
      p += cnt[static_cast<size_t>(w) * nb + b];
    }
  }
  for (int64_t b = 0; b < nb; b++) {
    int64_t p = bstart[b];
    for (int32_t w = 0; w < nwin; w++) {
      bpos[static_cast<size_t>(w) * nb + b] = p;
      p += cnt[static_cast<size_t>(w) * nb + b];
    }
  }
  // 4. Segment tables per window (non-empty segments, bucket order).
  std::vector<int64_t> wseg(nwin + 1, 0);
  std::vector<int32_t> seg_e;
  std::vector<int32_t> seg_d;
  for (int32_t w = 0; w < nwin; w++) {
    for (int64_t b = 0; b < nb; b++) {
      const size_t k = static_cast<size_t>(w) * nb + b;
      if (!cnt[k]) continue;
      const int64_t e = wpos[k] - wstart[w];
      seg_e.push_back(static_cast<int32_t>(e));
      seg_d.push_back(static_cast<int32_t>(bpos[k] - e));
    }
    wseg[w + 1] = static_cast<int64_t>(seg_e.size());
  }
  // 5. Fill both arrays, bucket by bucket (disjoint segments).
  std::vector<uint16_t> src(std::max<int64_t>(1, wstart[nwin]), 0);
  std::vector<uint16_t> brow_off(std::max<int64_t>(1, m), 0);
  parallel_dynamic(nb, nt, 16, [&](int64_t b) {
    for (int64_t r = brow[b]; r < brow[b + 1]; r++)
      for (int64_t e = off[r]; e < off[r + 1]; e++) {
        const int32_t u = nbr[e];
        const size_t k = static_cast<size_t>(u / kBWin) * nb + b;
        src[wpos[k]++] = static_cast<uint16_t>(u % kBWin);
        brow_off[bpos[k]++] = static_cast<uint16_t>(r - brow[b]);
      }
  });
  // 6. Units: a bucket whole, or a one-row bucket beyond kBEdges in parts
  // whose sums go to partial slots.
  std::vector<BlockUnit> units;
  std::vector<BlockSplit> splits;
  int32_t slots = 0;
  for (int64_t b = 0; b < nb; b++) {
    const int64_t e0 = bstart[b], e1 = bstart[b + 1];
    if (e1 - e0 <= kBEdges) {
      units.push_back({brow[b], brow[b + 1] - brow[b], -1, 0, e0, e1});
      continue;
    }
    const int32_t first = slots;
    for (int64_t e = e0; e < e1; e += kBEdges) units.push_back({brow[b], 1, slots++, 0, e, std::min(e1, e + kBEdges)});
    splits.push_back({brow[b], first, slots - first, 0});
  }
  out->src.swap(src);
  out->wstart.swap(wstart);
  out->wlen.swap(wlen);
  out->wseg.swap(wseg);
  out->seg_e.swap(seg_e);
  out->seg_d.swap(seg_d);
  out->brow_off.swap(brow_off);
  out->units.swap(units);
  out->splits.swap(splits);
  out->n_slots = slots;
  out->n_windows = nwin;
  out->n_buckets = nb;
  return DMM_OK;
}

namespace {

template <typename T>
int upload_vec(dmm_ctx* ctx, void** dst, const std::vector<T>& v, const char* what) {
  int rc = dev_alloc(ctx, dst, sizeof(T) * std::max<size_t>(1, v.size()), what);
  if (rc) return rc;
  if (!v.empty()) DMM_HIP_CHECK(ctx, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return DMM_OK;
}

}  // namespace

int pr_blocks_build(dmm_ctx* ctx) {
  const int64_t n = ctx->nloc;
  const int64_t m = ctx->m_in;
  std::vector<int64_t> off(n + 1);
  std::vector<int32_t> nbr(std::max<int64_t>(1, m));
  DMM_HIP_CHECK(ctx, hipMemcpy(off.data(), ctx->d_in_off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m > 0) DMM_HIP_CHECK(ctx, hipMemcpy(nbr.data(), ctx->d_in_nbr, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  BlockHost bh;
  std::string err;
  const int64_t held = ctx->dev_bytes;
  int rc = build_blocks_host(n, ctx->n, off.data(), nbr.data(), &bh, &err);
  if (rc) return set_err(ctx, rc, "%s", err.c_str());
  std::vector<int64_t>().swap(off);
  std::vector<int32_t>().swap(nbr);
  if ((rc = upload_vec(ctx, &ctx->d_bp_src, bh.src, "PR block sources")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_wstart, bh.wstart, "PR block window starts")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_wlen, bh.wlen, "PR block window lengths")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_wseg, bh.wseg, "PR block window segments")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_seg_e, bh.seg_e, "PR block segment starts")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_seg_d, bh.seg_d, "PR block segment destinations")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_brow, bh.brow_off, "PR block row offsets")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_units, bh.units, "PR block units")) ||
      (rc = upload_vec(ctx, &ctx->d_bp_splits, bh.splits, "PR block split rows")) ||
      // One spare slot past the m contributions takes the propagation
      // pass's stores from lanes past a window's end.
      (rc = dev_alloc(ctx, &ctx->d_bp_val, sizeof(float) * (m + 1), "PR block contributions")) ||
      (rc = dev_alloc(ctx, &ctx->d_bp_part, 16 * std::max<int64_t>(1, bh.n_slots), "PR block split partials"))) {
    ctx->bp_bytes = ctx->dev_bytes - held;  // what the failed build holds, freed with the layout
    return rc;
  }
  ctx->n_bp_windows = bh.n_windows;
  ctx->n_bp_units = static_cast<int64_t>(bh.units.size());
  ctx->n_bp_splits = static_cast<int64_t>(bh.splits.size());
  ctx->n_bp_buckets = bh.n_buckets;
  ctx->bp_bytes = ctx->dev_bytes - held;
  ctx->bp_ready = true;
  return DMM_OK;
}

namespace {

constexpr int kHostFxShift = 80;  // as the device's kFxShift

// Host restatement of the device's fx_of: c * 2^80 as a 128-bit two's
// complement integer, truncated below 2^-80.
__int128 host_fx_of(float c) {
  uint32_t bits;
  std::memcpy(&bits, &c, 4);
  const int e = static_cast<int>((bits >> 23) & 0xFFu);
  const unsigned __int128 m = (bits & 0x7FFFFFu) | 0x800000u;
  if (e == 0) return 0;
  const int sh = e - 150 + kHostFxShift;
  if (sh <= -24) return 0;
  const unsigned __int128 x = sh < 0 ? m >> (-sh) : m << sh;
  return (bits >> 31) ? -static_cast<__int128>(x) : static_cast<__int128>(x);
}

// Host restatement of fx_value: the top 64 significant bits with a sticky
// bit, one rounding to FP32, then the exact power-of-two scaling.
float host_fx_value(__int128 v) {
  const bool neg = v < 0;
  const unsigned __int128 a = neg ? static_cast<unsigned __int128>(-v) : static_cast<unsigned __int128>(v);
  const uint64_t hi = static_cast<uint64_t>(a >> 64), lo = static_cast<uint64_t>(a);
  float f;
  if (hi == 0) {
    f = std::ldexp(static_cast<float>(lo), -kHostFxShift);
  } else {
    const int lz = __builtin_clzll(hi);
    const uint64_t top = lz ? (hi << lz) | (lo >> (64 - lz)) : hi;
    const uint64_t rest = lz ? lo << lz : lo;
    f = std::ldexp(static_cast<float>(top | (rest != 0 ? 1ull : 0ull)), 64 - lz - kHostFxShift);
  }
  return neg ? -f : f;
}

}  // namespace

}  // namespace dmm

// Host check of the PE-block propagation layout (no GPU): builds it, checks
// every index the kernels use against its buffer and that every edge reaches
// its row once, and replays pass 1 and pass 2 with the kernels' arithmetic.
extern "C" int dmm_pr_blocks_check(int64_t n, int64_t n_gather, const int64_t* off, const int32_t* nbr,
                                   const float* contrib, float* incoming, int64_t* stats) {
  using namespace dmm;
  if (n < 0 || !off || (off[n] > 0 && !nbr) || !contrib || !incoming || !stats)
    return set_err(nullptr, DMM_ERR_INVALID, "dmm_pr_blocks_check: bad arguments");
  const int64_t m = off[n];
  for (int64_t e = 0; e < m; e++)
    if (nbr[e] < 0 || nbr[e] >= n_gather) return set_err(nullptr, DMM_ERR_GRAPH, "source outside [0, n_gather)");
  BlockHost bh;
  std::string err;
  int rc = build_blocks_host(n, n_gather, off, nbr, &bh, &err);
  if (rc) return set_err(nullptr, rc, "%s", err.c_str());
  auto bad = [](const char* what) { return set_err(nullptr, DMM_ERR_GRAPH, "block layout: %s", what); };
  const int32_t nwin = bh.n_windows;
  if (static_cast<int32_t>(bh.wstart.size()) != nwin + 1 || static_cast<int32_t>(bh.wseg.size()) != nwin + 1 ||
      static_cast<int64_t>(bh.brow_off.size()) < std::max<int64_t>(1, m) || bh.n_buckets > kBSegMax)
    return bad("sizes");
  // Pass 1, window by window, as k_bprop walks it.
  std::vector<float> val(std::max<int64_t>(1, m), 0.0f);
  std::vector<char> hit(std::max<int64_t>(1, m), 0);
  for (int32_t w = 0; w < nwin; w++) {
    const int64_t base = int64_t(w) * kBWin;
    const int64_t cnt = std::min<int64_t>(kBWin, n_gather - base);
    const int64_t s0 = bh.wseg[w];
    const int ns = static_cast<int>(bh.wseg[w + 1] - s0);
    const int len = bh.wlen[w];
    if (ns < 0 || ns > kBSegMax || bh.wstart[w] % kBPer || bh.wstart[w] + len > bh.wstart[w + 1]) return bad("window");
    if (ns > 0 && bh.seg_e[s0] != 0) return bad("first segment");
    for (int k = 1; k < ns; k++)
      if (bh.seg_e[s0 + k] <= bh.seg_e[s0 + k - 1] || bh.seg_e[s0 + k] >= len) return bad("segment starts");
    for (int i0 = 0; i0 < len; i0 += kBPer) {
      int lo = 0, hi = ns;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bh.seg_e[s0 + mid] <= i0) lo = mid; else hi = mid;
      }
      int next = lo + 1 < ns ? bh.seg_e[s0 + lo + 1] : len;
      int64_t dst = int64_t(bh.seg_d[s0 + lo]) + i0;
      for (int k = 0; k < kBPer && i0 + k < len; k++) {
        const int i = i0 + k;
        if (i == next) {
          lo++;
          next = lo + 1 < ns ? bh.seg_e[s0 + lo + 1] : len;
          dst = int64_t(bh.seg_d[s0 + lo]) + i;
        }
        const uint16_t s = bh.src[bh.wstart[w] + i];
        if (s >= cnt) return bad("source offset");
        if (dst < 0 || dst >= m || hit[dst]) return bad("destination");
        hit[dst] = 1;
        val[dst] = contrib[base + s];
        dst++;
      }
    }
  }
  for (int64_t e = 0; e < m; e++)
    if (!hit[e]) return bad("destination never written");
  // Pass 2, unit by unit, and the split rows; every row finishes once and
  // every unit holds exactly its rows' in-edges.
  std::vector<__int128> part(std::max(1, bh.n_slots), 0);
  std::vector<int64_t> slot_edges(std::max(1, bh.n_slots), 0);
  std::vector<char> done(std::max<int64_t>(1, n), 0);
  int64_t next_e = 0;
  for (const BlockUnit& u : bh.units) {
    if (u.e0 != next_e || u.e1 < u.e0 || u.nrows < 1 || u.nrows > kBRows || u.e1 - u.e0 > kBEdges ||
        u.row0 < 0 || int64_t(u.row0) + u.nrows > n)
      return bad("unit");
    next_e = u.e1;
    if (u.slot >= 0) {
      if (u.nrows != 1 || u.slot >= bh.n_slots) return bad("split unit");
      __int128 s = 0;
      for (int64_t e = u.e0; e < u.e1; e++) {
        if (bh.brow_off[e] != 0) return bad("split unit row");
        s += host_fx_of(val[e]);
      }
      part[u.slot] = s;
      slot_edges[u.slot] = u.e1 - u.e0;
      continue;
    }
    if (off[u.row0 + u.nrows] - off[u.row0] != u.e1 - u.e0) return bad("unit edges differ from its rows' in-edges");
    std::vector<__int128> acc(u.nrows, 0);
    for (int64_t e = u.e0; e < u.e1; e++) {
      if (bh.brow_off[e] >= u.nrows) return bad("row offset");
      acc[bh.brow_off[e]] += host_fx_of(val[e]);
    }
    for (int32_t r = 0; r < u.nrows; r++) {
      incoming[u.row0 + r] = host_fx_value(acc[r]);
      done[u.row0 + r]++;
    }
  }
  if (next_e != m) return bad("units do not cover the edges");
  for (const BlockSplit& sp : bh.splits) {
    if (sp.row < 0 || sp.row >= n || sp.first < 0 || sp.count < 1 || sp.first + sp.count > bh.n_slots)
      return bad("split row");
    __int128 s = 0;
    int64_t edges = 0;
    for (int j = 0; j < sp.count; j++) {
      s += part[sp.first + j];
      edges += slot_edges[sp.first + j];
    }
    if (edges != off[sp.row + 1] - off[sp.row]) return bad("split row edges");
    incoming[sp.row] = host_fx_value(s);
    done[sp.row]++;
  }
  for (int64_t r = 0; r < n; r++)
    if (done[r] != 1) return bad("row not finished exactly once");
  stats[0] = nwin;
  stats[1] = bh.n_buckets;
  stats[2] = static_cast<int64_t>(bh.units.size());
  stats[3] = static_cast<int64_t>(bh.splits.size());
  stats[4] = static_cast<int64_t>(bh.seg_e.size());
  stats[5] = bh.n_slots;
  return DMM_OK;
}
