// Host-side construction of the window-tiled compensated PageRank layout.
//
// Input: the degree-ordered in-CSR already on the device.  Every row's
// in-neighbours are sorted and cut by source window (window = the
// `window` consecutive source IDs whose contributions one workgroup holds
// in LDS).  A row's in-edges inside one window form a run; every run of
// every window, hub window included, is summed by the tile kernel, so no
// contribution is gathered from HBM.
//
// Tile layout (window-major, rows ascending inside a window):
//   chunk    = 64 lanes x kLaneEdges edge slots, all of one window, holding
//              at most kChunkPieces pieces;
//   piece    = the part of a run inside one chunk (a run that crosses a
//              chunk boundary is cut there); its sum is one partial;
//   per lane = kLaneEdges 16-bit source offsets and an end mask (bit j:
//              slot j is the last edge of a piece) with the count of
//              filled slots.
// Partials are numbered in chunk order, so a chunk writes one contiguous
// range of them and every partial's slot follows from the chunk's first
// slot and the piece's rank in lane order.
//
// Row blocks: rows are grouped into blocks whose partials fit one
// workgroup's LDS.  A block's partials in one window are contiguous (a
// segment); the row kernel gathers the block's segments, in window order,
// into LDS (the block's region) and merges every row's partials through a
// list of region offsets kept in window order, so iterations are
// bit-deterministic.  Rows with many runs or many in-edges form blocks of
// their own.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../host/parallel.h"
#include "pr.h"

namespace dmm {

namespace {

struct RunRec {
  int32_t row;
  int32_t len;
  int64_t begin;  // first edge in the sorted neighbour array
};

struct PieceRec {
  int32_t row;
  int32_t block;
};

struct WinPack {
  std::vector<uint16_t> src;     // kChunkEdges per chunk
  std::vector<uint32_t> mask;    // per lane: end mask | filled slots << 16
  std::vector<int32_t> cpieces;  // pieces per chunk
  std::vector<PieceRec> pieces;  // in chunk order
};

template <typename T>
int upload_vec(dmm_ctx* ctx, void** dst, const std::vector<T>& v, const char* what) {
  int rc = dev_alloc(ctx, dst, sizeof(T) * std::max<size_t>(1, v.size()), what);
  if (rc) return rc;
  if (!v.empty()) DMM_HIP_CHECK(ctx, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return DMM_OK;
}

}  // namespace

// Builds the tile layout on the host from an in-CSR of n rows whose
// sources lie in [0, n_gather); rows at and above zb have no in-edges.
// nbr is sorted row by row in place.  Returns DMM_OK or an error code
// (message in *err).
int build_tiles_host(int64_t n, int64_t zb, int64_t n_gather, const int64_t* off, int32_t* nbr, int64_t W,
                     TileHost* out, std::string* err) {
  const int nt = resolve_threads(0);
  if (W != 16384 && W != 32768) {
    *err = "PageRank tile window is not 16384 or 32768";
    return DMM_ERR_INVALID;
  }
  if (zb < 0 || zb > n) {
    *err = "bad in-degree-0 tail";
    return DMM_ERR_INVALID;
  }
  const int32_t nwin = static_cast<int32_t>(std::max<int64_t>(1, (n_gather + W - 1) / W));
  // 1. Rows sorted; runs per row.
  std::vector<int32_t> runs(std::max<int64_t>(1, zb), 0);
  parallel_dynamic(zb, nt, 4096, [&](int64_t r) {
    int32_t* b = nbr + off[r];
    int32_t* e = nbr + off[r + 1];
    std::sort(b, e);
    int32_t c = 0;
    for (int32_t* p = b; p < e;) {
      const int64_t w = *p / W;
      while (p < e && *p / W == w) p++;
      c++;
    }
    runs[r] = c;
  });
  // 2. Row blocks over [0, zb).
  std::vector<int32_t> brow{0};
  {
    int64_t cur = 0;
    for (int64_t r = 0; r < zb; r++) {
      const bool solo = runs[r] >= kSoloRuns || off[r + 1] - off[r] >= kSoloEdges;
      if (solo) {
        if (brow.back() != r) brow.push_back(static_cast<int32_t>(r));
        brow.push_back(static_cast<int32_t>(r + 1));
        cur = 0;
        continue;
      }
      if (cur + runs[r] > kBlockRuns && brow.back() != r) {
        brow.push_back(static_cast<int32_t>(r));
        cur = 0;
      }
      cur += runs[r];
    }
    if (brow.back() != zb) brow.push_back(static_cast<int32_t>(zb));
  }
  const int64_t nb = static_cast<int64_t>(brow.size()) - 1;
  std::vector<int32_t> block_of(std::max<int64_t>(1, zb));
  for (int64_t b = 0; b < nb; b++)
    for (int64_t r = brow[b]; r < brow[b + 1]; r++) block_of[r] = static_cast<int32_t>(b);
  // 3. Runs per window, rows ascending (thread ranges are row ranges).
  std::vector<std::vector<std::vector<RunRec>>> recs(nt, std::vector<std::vector<RunRec>>(nwin));
  parallel_ranges(zb, nt, [&](int t, int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; r++) {
      const int32_t* b = nbr + off[r];
      const int32_t* e = nbr + off[r + 1];
      for (const int32_t* p = b; p < e;) {
        const int64_t w = *p / W;
        const int32_t* q = p;
        while (q < e && *q / W == w) q++;
        recs[t][w].push_back({static_cast<int32_t>(r), static_cast<int32_t>(q - p), p - nbr});
        p = q;
      }
    }
  });
  // 4. Pack every window into chunks.
  std::vector<WinPack> packs(nwin);
  parallel_dynamic(nwin, nt, 1, [&](int64_t w) {
    WinPack& o = packs[w];
    const int64_t base = w * W;
    int64_t ch = -1;
    int lane = 64, slot = 0;
    int64_t last_lane = -1;
    int last_slot = -1;
    auto new_chunk = [&]() {
      ch++;
      lane = 0;
      slot = 0;
      o.src.resize((ch + 1) * kChunkEdges, 0);
      o.mask.resize((ch + 1) * 64, 0u);
      o.cpieces.push_back(0);
    };
    auto start_piece = [&](int32_t row, int32_t blk) {
      o.pieces.push_back({row, blk});
      o.cpieces[ch]++;
    };
    for (int t = 0; t < nt; t++)
      for (const RunRec& rr : recs[t][w]) {
        const int32_t blk = block_of[rr.row];
        if (slot == kLaneEdges) {
          lane++;
          slot = 0;
        }
        if (lane >= 64 || o.cpieces[ch] == kChunkPieces) new_chunk();
        start_piece(rr.row, blk);
        for (int32_t k = 0; k < rr.len; k++) {
          if (slot == kLaneEdges) {
            lane++;
            slot = 0;
            if (lane == 64) {  // the run crosses a chunk boundary: cut it
              o.mask[last_lane] |= 1u << last_slot;
              new_chunk();
              start_piece(rr.row, blk);
            }
          }
          const int64_t li = ch * 64 + lane;
          o.src[li * kLaneEdges + slot] = static_cast<uint16_t>(nbr[rr.begin + k] - base);
          o.mask[li] = (o.mask[li] & 0xFFFFu) | (static_cast<uint32_t>(slot + 1) << 16);
          last_lane = li;
          last_slot = slot;
          slot++;
        }
        o.mask[last_lane] |= 1u << last_slot;  // the piece ends here
      }
  });
  recs.clear();
  recs.shrink_to_fit();
  // 5. Chunks in window order and their first partials.
  std::vector<int64_t> wstart(nwin + 1, 0);
  for (int32_t w = 0; w < nwin; w++) wstart[w + 1] = wstart[w] + static_cast<int64_t>(packs[w].cpieces.size());
  const int64_t nchunks = wstart[nwin];
  std::vector<int32_t> cbase(nchunks + 1, 0);
  {
    int64_t g = 0;
    for (int32_t w = 0; w < nwin; w++)
      for (size_t k = 0; k < packs[w].cpieces.size(); k++) {
        cbase[wstart[w] + k] = static_cast<int32_t>(g);
        g += packs[w].cpieces[k];
        if (g >= (int64_t(1) << 31)) {
          *err = "PageRank tile partials exceed 2^31";
          return DMM_ERR_CAPACITY;
        }
      }
    cbase[nchunks] = static_cast<int32_t>(g);
  }
  const int64_t npieces = cbase[nchunks];
  // 6. Segments per block, in window order; region offsets; row lists.
  std::vector<std::vector<int32_t>> bs_start(nb), bs_off(nb);
  std::vector<int64_t> btot(nb, 0);
  std::vector<int32_t> rcnt(std::max<int64_t>(1, zb), 0);
  std::vector<char> multi(nb, 0);
  for (int64_t b = 0; b < nb; b++) multi[b] = brow[b + 1] - brow[b] > 1;
  std::vector<int32_t> rpos(npieces > 0 ? npieces : 1);  // region offset of every partial
  for (int32_t w = 0; w < nwin; w++) {
    const std::vector<PieceRec>& pc = packs[w].pieces;
    const int64_t g0 = cbase[wstart[w]];
    for (size_t i = 0; i < pc.size();) {
      const int32_t b = pc[i].block;
      size_t j = i;
      while (j < pc.size() && pc[j].block == b) j++;
      bs_start[b].push_back(static_cast<int32_t>(g0 + i));
      bs_off[b].push_back(static_cast<int32_t>(btot[b]));
      for (size_t k = i; k < j; k++) {
        rpos[g0 + k] = static_cast<int32_t>(btot[b]++);
        if (multi[b]) rcnt[pc[k].row]++;
      }
      i = j;
    }
  }
  for (int64_t b = 0; b < nb; b++)
    if (btot[b] > kRowSlots) {
      *err = "a row block has more PageRank tile partials than the row kernel stages";
      return DMM_ERR_CAPACITY;
    }
  std::vector<int64_t> bseg(nb + 1, 0);
  std::vector<int32_t> bsize(std::max<int64_t>(1, nb), 0);
  for (int64_t b = 0; b < nb; b++) {
    bseg[b + 1] = bseg[b] + static_cast<int64_t>(bs_start[b].size());
    bsize[b] = static_cast<int32_t>(btot[b]);
  }
  std::vector<int32_t> istart(zb + 1, 0);
  for (int64_t r = 0; r < zb; r++) {
    const int64_t next = int64_t(istart[r]) + rcnt[r];
    if (next >= (int64_t(1) << 31)) {
      *err = "PageRank tile row lists exceed 2^31 entries";
      return DMM_ERR_CAPACITY;
    }
    istart[r + 1] = static_cast<int32_t>(next);
  }
  std::vector<int32_t> seg_start(std::max<int64_t>(1, bseg[nb])), seg_off(std::max<int64_t>(1, bseg[nb]));
  for (int64_t b = 0; b < nb; b++) {
    std::copy(bs_start[b].begin(), bs_start[b].end(), seg_start.begin() + bseg[b]);
    std::copy(bs_off[b].begin(), bs_off[b].end(), seg_off.begin() + bseg[b]);
  }
  bs_start.clear();
  bs_off.clear();
  std::vector<uint16_t> idx(std::max<int64_t>(1, istart[zb]));
  {
    std::vector<int32_t> fill(istart.begin(), istart.end() - 1);
    for (int32_t w = 0; w < nwin; w++) {
      const std::vector<PieceRec>& pc = packs[w].pieces;
      const int64_t g0 = cbase[wstart[w]];
      for (size_t i = 0; i < pc.size(); i++)
        if (multi[pc[i].block]) idx[fill[pc[i].row]++] = static_cast<uint16_t>(rpos[g0 + i]);
    }
  }
  // 7. Chunk arrays.
  std::vector<int32_t> cwin(std::max<int64_t>(1, nchunks));
  std::vector<uint16_t> src(std::max<int64_t>(1, nchunks * kChunkEdges));
  std::vector<uint32_t> meta(std::max<int64_t>(1, nchunks * 64));
  parallel_dynamic(nwin, nt, 1, [&](int64_t w) {
    const WinPack& o = packs[w];
    const int64_t c0 = wstart[w], nc = wstart[w + 1] - c0;
    for (int64_t c = 0; c < nc; c++) cwin[c0 + c] = static_cast<int32_t>(w);
    std::copy(o.src.begin(), o.src.end(), src.begin() + c0 * kChunkEdges);
    std::copy(o.mask.begin(), o.mask.end(), meta.begin() + c0 * 64);
  });
  int64_t filled = 0;
  for (int32_t w = 0; w < nwin; w++)
    for (uint32_t mk : packs[w].mask) filled += mk >> 16;
  packs.clear();
  packs.shrink_to_fit();
  out->cwin.swap(cwin);
  out->wstart.swap(wstart);
  out->cbase.swap(cbase);
  out->meta.swap(meta);
  out->src.swap(src);
  out->brow.swap(brow);
  out->bseg.swap(bseg);
  out->bsize.swap(bsize);
  out->seg_start.swap(seg_start);
  out->seg_off.swap(seg_off);
  out->istart.swap(istart);
  out->idx.swap(idx);
  out->n_pieces = npieces;
  out->slots_filled = filled;
  out->window = W;
  out->n_windows = nwin;
  return DMM_OK;
}

int pr_tiles_build(dmm_ctx* ctx) {
  const int64_t n = ctx->nloc;
  const int64_t m = ctx->m_in;
  std::vector<int64_t> off(n + 1);
  std::vector<int32_t> nbr(std::max<int64_t>(1, m));
  DMM_HIP_CHECK(ctx, hipMemcpy(off.data(), ctx->d_in_off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m > 0) DMM_HIP_CHECK(ctx, hipMemcpy(nbr.data(), ctx->d_in_nbr, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  TileHost th;
  std::string err;
  const int64_t held = ctx->dev_bytes;
  int rc = build_tiles_host(n, ctx->pr_zero_begin, ctx->n, off.data(), nbr.data(), ctx->tuning.pr_window, &th, &err);
  if (rc) return set_err(ctx, rc, "%s", err.c_str());
  std::vector<int64_t>().swap(off);
  std::vector<int32_t>().swap(nbr);
  if ((rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_cwin), th.cwin, "PR tile chunk windows")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_wstart), th.wstart, "PR tile window starts")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_cbase), th.cbase, "PR tile chunk partials")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_meta), th.meta, "PR tile lanes")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_src), th.src, "PR tile sources")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_brow), th.brow, "PR row blocks")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_bseg), th.bseg, "PR row block segments")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_bsize), th.bsize, "PR row block sizes")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_seg_start), th.seg_start, "PR segments")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_seg_off), th.seg_off, "PR segments")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_istart), th.istart, "PR row lists")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_wt_idx), th.idx, "PR row piece lists")) ||
      (rc = dev_alloc(ctx, &ctx->d_wt_part, sizeof(float) * std::max<int64_t>(1, th.n_pieces), "PR tile partials"))) {
    ctx->tiles_bytes = ctx->dev_bytes - held;  // what the failed build holds, freed with the layout
    return rc;
  }
  ctx->n_chunks = static_cast<int64_t>(th.wstart.back());
  ctx->n_windows = th.n_windows;
  ctx->n_pieces = th.n_pieces;
  ctx->n_row_blocks = static_cast<int64_t>(th.brow.size()) - 1;
  ctx->tile_slots_filled = th.slots_filled;
  ctx->tiles_window = th.window;
  ctx->tiles_bytes = ctx->dev_bytes - held;
  ctx->tiles_ready = true;
  return DMM_OK;
}

namespace {

// Host restatement of the compensated sum (the device's Comp): TwoSum of
// the running sum and the new term, errors accumulated separately.
struct HostComp {
  float s = 0.0f, c = 0.0f;
  void merge(float s2, float c2) {
    const float t = s + s2;
    const float bb = t - s;
    const float e = (s - (t - bb)) + (s2 - bb);
    s = t;
    c = (c + c2) + e;
  }
  void add(float x) {  // the device's Comp::add
    const float t = s + x;
    const float bb = t - s;
    const float e = (s - (t - bb)) + (x - bb);
    s = t;
    c += e;
  }
  float value() const { return s + c; }
};

// The device's fixed butterfly over 64 lanes (FusedAcc<float>::reduce<64>).
HostComp butterfly64(HostComp* v) {
  for (int d = 32; d >= 1; d >>= 1) {
    HostComp nv[64];
    for (int l = 0; l < 64; l++) {
      const HostComp& p = v[l ^ d];
      if (l & d) {
        HostComp o = p;
        o.merge(v[l].s, v[l].c);
        nv[l] = o;
      } else {
        nv[l] = v[l];
        nv[l].merge(p.s, p.c);
      }
    }
    for (int l = 0; l < 64; l++) v[l] = nv[l];
  }
  return v[0];
}

// The device's segmented inclusive scan (seg_scan in pr.hip), step for
// step, so the FP64 sums round identically.
void host_seg_scan(const double* tail, const bool* has_end, double* sc) {
  int seg[64];
  int h = -1;
  for (int l = 0; l < 64; l++) {
    if (has_end[l]) h = l;
    seg[l] = h < 0 ? 0 : h;
    sc[l] = tail[l];
  }
  double nv[64];
  for (int d = 1; d <= 8; d <<= 1) {
    for (int l = 0; l < 64; l++) nv[l] = ((l & 15) >= d && l - d >= seg[l]) ? sc[l] + sc[l - d] : sc[l];
    std::copy(nv, nv + 64, sc);
  }
  for (int l = 0; l < 64; l++)
    nv[l] = (((l >> 4) & 1) && (l & ~15) - 1 >= seg[l]) ? sc[l] + sc[(l & ~15) - 1] : sc[l];
  std::copy(nv, nv + 64, sc);
  for (int l = 0; l < 64; l++) nv[l] = (l >= 32 && 31 >= seg[l]) ? sc[l] + sc[31] : sc[l];
  std::copy(nv, nv + 64, sc);
}

}  // namespace

}  // namespace dmm

// Host check of the tile layout (no GPU): builds it, verifies every index
// the kernels use against its buffer, and replays one iteration's sums with
// the kernels' arithmetic (k_tile, k_rows) on the host.
extern "C" int dmm_pr_tiles_check(int64_t n, int64_t zb, int64_t n_gather, const int64_t* off, const int32_t* nbr,
                                  int64_t window, const float* contrib, float* incoming, int64_t* stats) {
  using namespace dmm;
  if (n < 0 || !off || (off[n] > 0 && !nbr) || !contrib || !incoming || !stats)
    return set_err(nullptr, DMM_ERR_INVALID, "dmm_pr_tiles_check: bad arguments");
  const int64_t m = off[n];
  for (int64_t e = 0; e < m; e++)
    if (nbr[e] < 0 || nbr[e] >= n_gather) return set_err(nullptr, DMM_ERR_GRAPH, "source outside [0, n_gather)");
  std::vector<int32_t> sorted(nbr, nbr + m);
  TileHost th;
  std::string err;
  int rc = build_tiles_host(n, zb, n_gather, off, sorted.data(), window, &th, &err);
  if (rc) return set_err(nullptr, rc, "%s", err.c_str());
  const int64_t nchunks = th.wstart.back(), np = th.n_pieces, nb = static_cast<int64_t>(th.brow.size()) - 1;
  auto bad = [](const char* what) { return set_err(nullptr, DMM_ERR_GRAPH, "tile layout: %s", what); };
  if (static_cast<int64_t>(th.cwin.size()) < nchunks || static_cast<int64_t>(th.meta.size()) < nchunks * 64 ||
      static_cast<int64_t>(th.src.size()) < nchunks * kChunkEdges ||
      static_cast<int64_t>(th.cbase.size()) != nchunks + 1 || th.cbase[nchunks] != np || th.brow[0] != 0 ||
      th.brow.back() != zb)
    return bad("sizes");
  std::vector<char> written(std::max<int64_t>(1, np), 0);
  std::vector<float> px(std::max<int64_t>(1, np));
  int64_t edges = 0;
  // k_tile, chunk by chunk.
  for (int32_t w = 0; w < th.n_windows; w++) {
    const int64_t base = int64_t(w) * window;
    const int64_t cnt_w = std::min<int64_t>(window, n_gather - base);
    for (int64_t c = th.wstart[w]; c < th.wstart[w + 1]; c++) {
      if (th.cwin[c] != w) return bad("chunk window");
      double tail[64];
      bool has_end[64];
      float x[64][kLaneEdges];
      int total = 0;
      for (int l = 0; l < 64; l++) {
        const uint32_t mt = th.meta[c * 64 + l];
        const uint32_t ends = mt & 0xFFFFu;
        const int cnt = static_cast<int>(mt >> 16);
        if (cnt > kLaneEdges || (cnt < kLaneEdges && (ends >> cnt) != 0)) return bad("lane count / end mask");
        edges += cnt;
        total += __builtin_popcount(ends);
        // The lane's first pass (tile_chunk in pr.hip): its sum after its
        // last piece end.
        double acc = 0.0;
        for (int j = 0; j < kLaneEdges; j++) {
          const uint16_t s = th.src[c * kChunkEdges + l * kLaneEdges + j];
          // Unfilled slots are summed unmasked by the tile kernel: offset 0.
          if (j < cnt ? s >= cnt_w : s != 0) return bad("source offset");
          x[l][j] = j < cnt ? contrib[base + s] : 0.0f;
          acc += static_cast<double>(x[l][j]);
          if ((ends >> j) & 1u) acc = 0.0;
        }
        tail[l] = acc;
        has_end[l] = ends != 0;
      }
      if (total > kChunkPieces || total != th.cbase[c + 1] - th.cbase[c]) return bad("chunk piece count");
      double sc[64];
      host_seg_scan(tail, has_end, sc);
      // The lane's second pass: its pieces, the first one from the carry,
      // numbered in lane order from the chunk's first partial.
      int64_t slot = th.cbase[c];
      for (int l = 0; l < 64; l++) {
        const uint32_t ends = th.meta[c * 64 + l] & 0xFFFFu;
        double a = l == 0 ? 0.0 : sc[l - 1];
        for (int j = 0; j < kLaneEdges; j++) {
          a += static_cast<double>(x[l][j]);
          if ((ends >> j) & 1u) {
            if (slot >= th.cbase[c + 1] || written[slot]) return bad("output slot");
            written[slot] = 1;
            px[slot] = static_cast<float>(a);
            slot++;
            a = 0.0;
          }
        }
      }
      // Every piece ends inside its chunk: the last filled slot ends one,
      // and every lane before it is full or ends a piece at its last slot.
      int last = -1;
      for (int l = 0; l < 64; l++)
        if ((th.meta[c * 64 + l] >> 16) != 0) last = l;
      if (last < 0) return bad("empty chunk");
      for (int l = 0; l <= last; l++) {
        const uint32_t mt = th.meta[c * 64 + l];
        const uint32_t cl = mt >> 16;
        if (cl == 0 || (cl < kLaneEdges && !((mt >> (cl - 1)) & 1u))) return bad("short lane");
      }
      const uint32_t cl = th.meta[c * 64 + last] >> 16;
      if (!((th.meta[c * 64 + last] >> (cl - 1)) & 1u)) return bad("piece open at chunk end");
    }
  }
  if (edges != m) return bad("edge count");
  for (int64_t p = 0; p < np; p++)
    if (!written[p]) return bad("slot never written");
  // k_rows: each block's region gathered from its segments, window order.
  int64_t max_block = 0;
  std::vector<float> rx;
  for (int64_t b = 0; b < nb; b++) {
    const int32_t r0 = th.brow[b], r1 = th.brow[b + 1];
    if (r1 <= r0) return bad("row block bounds");
    rx.clear();
    for (int64_t s = th.bseg[b]; s < th.bseg[b + 1]; s++) {
      const int64_t len = (s + 1 < th.bseg[b + 1] ? th.seg_off[s + 1] : th.bsize[b]) - th.seg_off[s];
      if (th.seg_off[s] != static_cast<int64_t>(rx.size()) || th.seg_start[s] < 0 || len <= 0 ||
          th.seg_start[s] + len > np)
        return bad("segment bounds");
      for (int64_t k = 0; k < len; k++) rx.push_back(px[th.seg_start[s] + k]);
    }
    if (static_cast<int64_t>(rx.size()) != th.bsize[b]) return bad("block size");
    const int64_t rn = static_cast<int64_t>(rx.size());
    if (rn > kRowSlots) return bad("block region beyond the row kernel's LDS");
    max_block = std::max(max_block, rn);
    if (r1 - r0 == 1) {
      HostComp lanes[kRowThreads];
      for (int t = 0; t < kRowThreads; t++)
        for (int64_t p = t; p < rn; p += kRowThreads) lanes[t].add(rx[p]);
      HostComp wv[kRowThreads / 64];
      for (int w = 0; w < kRowThreads / 64; w++) wv[w] = butterfly64(lanes + 64 * w);
      HostComp tot = wv[0];
      for (int w = 1; w < kRowThreads / 64; w++) tot.merge(wv[w].s, wv[w].c);
      incoming[r0] = tot.value();
      continue;
    }
    for (int32_t r = r0; r < r1; r++) {
      HostComp acc;
      if (th.istart[r + 1] < th.istart[r] || th.istart[r + 1] > static_cast<int64_t>(th.idx.size()))
        return bad("row list bounds");
      for (int64_t k = th.istart[r]; k < th.istart[r + 1]; k++) {
        const int64_t q = th.idx[k];
        if (q >= rn) return bad("partial offset beyond its block");
        acc.add(rx[q]);
      }
      incoming[r] = acc.value();
    }
  }
  for (int64_t r = zb; r < n; r++) incoming[r] = 0.0f;
  stats[0] = nchunks;
  stats[1] = np;
  stats[2] = nb;
  stats[3] = th.slots_filled;
  stats[4] = th.n_windows;
  stats[5] = max_block;
  return DMM_OK;
}
