// Host-side construction of the window-tiled compensated PageRank layout.
//
// Input: the degree-ordered in-CSR already on the device.  For each row the
// in-neighbours are sorted and cut by source window (kTileWindow vertices):
//   window 0                 -> the row's list (LDS hub copy in the row
//                               kernel)
//   run of >= kTileMinRun    -> pieces of <= kTilePiece edges, summed by the
//   in a window w >= 1          tile kernel from an LDS copy of window w,
//   holding >= kTileMin-        when the window's long runs hold at least
//   WindowEdges such edges      kTileMinWindowEdges edges
//   other runs               -> the row's list (gathered from HBM)
// Pieces are laid out window-major (rows ascending inside a window) and
// packed into 64-lane chunks without crossing a chunk; a piece's partial
// slot is its index in that order.  Each row keeps the slots of its pieces
// in window order, so the row kernel merges them in a fixed order and every
// iteration is bit-deterministic.
#include <algorithm>
#include <vector>

#include "../host/parallel.h"
#include "pr.h"

namespace dmm {

namespace {

struct PieceRec {
  int32_t row;
  int32_t len;
  int64_t begin;  // first edge in the sorted neighbour array
};

template <typename T>
int upload_vec(dmm_ctx* ctx, void** dst, const std::vector<T>& v, const char* what) {
  int rc = dev_alloc(ctx, dst, sizeof(T) * std::max<size_t>(1, v.size()), what);
  if (rc) return rc;
  if (!v.empty()) DMM_HIP_CHECK(ctx, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return DMM_OK;
}

}  // namespace

int pr_tiles_build(dmm_ctx* ctx) {
  const int64_t n = ctx->nloc;
  const int64_t m = ctx->m_in;
  const int nt = resolve_threads(0);
  // 1. The device CSR back on the host, rows sorted.
  std::vector<int64_t> off(n + 1);
  std::vector<int32_t> nbr(std::max<int64_t>(1, m));
  DMM_HIP_CHECK(ctx, hipMemcpy(off.data(), ctx->d_in_off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m > 0) DMM_HIP_CHECK(ctx, hipMemcpy(nbr.data(), ctx->d_in_nbr, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  const int32_t nwin = static_cast<int32_t>((ctx->n + kTileWindow - 1) / kTileWindow);
  // Sort every row; count per window the edges of runs long enough to tile.
  std::vector<std::vector<int64_t>> wlong(nt, std::vector<int64_t>(nwin, 0));
  parallel_ranges(n, nt, [&](int t, int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; r++) {
      int32_t* b = nbr.data() + off[r];
      int32_t* e = nbr.data() + off[r + 1];
      std::sort(b, e);
      for (int32_t* p = b; p < e;) {
        const int32_t w = *p / kTileWindow;
        int32_t* q = p;
        while (q < e && *q / kTileWindow == w) q++;
        if (w > 0 && q - p >= kTileMinRun) wlong[t][w] += q - p;
        p = q;
      }
    }
  });
  std::vector<char> tiled(nwin, 0);
  for (int32_t w = 1; w < nwin; w++) {
    int64_t s = 0;
    for (int t = 0; t < nt; t++) s += wlong[t][w];
    tiled[w] = s >= kTileMinWindowEdges;
  }
  std::vector<int64_t> lcnt(n + 1, 0), pcnt(n + 1, 0);
  parallel_dynamic(n, nt, 1024, [&](int64_t r) {
    const int32_t* b = nbr.data() + off[r];
    const int32_t* e = nbr.data() + off[r + 1];
    int64_t lc = 0, pc = 0;
    for (const int32_t* p = b; p < e;) {
      const int32_t w = *p / kTileWindow;
      const int32_t* q = p;
      while (q < e && *q / kTileWindow == w) q++;
      const int64_t len = q - p;
      if (!tiled[w] || len < kTileMinRun)
        lc += len;
      else
        pc += (len + kTilePiece - 1) / kTilePiece;
      p = q;
    }
    lcnt[r] = lc;
    pcnt[r] = pc;
  });
  // 2. Row offsets of the lists and of the piece slots.
  std::vector<int64_t> loff(n + 1, 0), poff(n + 1, 0);
  for (int64_t r = 0; r < n; r++) {
    loff[r + 1] = loff[r] + lcnt[r];
    poff[r + 1] = poff[r] + pcnt[r];
  }
  const int64_t nlist = loff[n], npieces = poff[n];
  // 3. Lists, and per thread and window the pieces in row order.
  std::vector<int32_t> lnbr(std::max<int64_t>(1, nlist));
  std::vector<std::vector<std::vector<PieceRec>>> recs(nt, std::vector<std::vector<PieceRec>>(nwin));
  parallel_ranges(n, nt, [&](int t, int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; r++) {
      const int32_t* b = nbr.data() + off[r];
      const int32_t* e = nbr.data() + off[r + 1];
      int64_t lo = loff[r];
      for (const int32_t* p = b; p < e;) {
        const int32_t w = *p / kTileWindow;
        const int32_t* q = p;
        while (q < e && *q / kTileWindow == w) q++;
        const int64_t len = q - p;
        if (!tiled[w] || len < kTileMinRun) {
          for (const int32_t* x = p; x < q; x++) lnbr[lo++] = *x;
        } else {
          for (int64_t s = 0; s < len; s += kTilePiece)
            recs[t][w].push_back({static_cast<int32_t>(r), static_cast<int32_t>(std::min<int64_t>(kTilePiece, len - s)),
                                  (p - nbr.data()) + s});
        }
        p = q;
      }
    }
  });
  // 4. Window-major slots and chunk counts.
  std::vector<int64_t> wpieces(nwin + 1, 0), wchunks(nwin + 1, 0);
  parallel_dynamic(nwin, nt, 1, [&](int64_t w) {
    int64_t pieces = 0, chunks = 0, fill = kTilePiece;
    for (int t = 0; t < nt; t++)
      for (const PieceRec& pr : recs[t][w]) {
        pieces++;
        if (fill + pr.len > kTilePiece) {
          chunks++;
          fill = 0;
        }
        fill += pr.len;
      }
    wpieces[w + 1] = pieces;
    wchunks[w + 1] = chunks;
  });
  for (int32_t w = 0; w < nwin; w++) {
    wpieces[w + 1] += wpieces[w];
    wchunks[w + 1] += wchunks[w];
  }
  const int64_t nchunks = wchunks[nwin];
  // 5. Chunks, source offsets and each row's slots (window order).
  std::vector<TileChunk> chunks(std::max<int64_t>(1, nchunks));
  std::vector<uint16_t> src(std::max<int64_t>(1, nchunks * 64), 0xFFFF);
  std::vector<int32_t> pslot(std::max<int64_t>(1, npieces));
  std::vector<int64_t> pnext(poff.begin(), poff.end() - 1);  // next free slot entry per row
  // Rows' slot entries must be filled in window order; windows are
  // processed in order, each row's entries appended as its pieces appear.
  for (int32_t w = 0; w < nwin; w++) {
    int64_t c = wchunks[w] - 1;
    int fill = kTilePiece;
    int64_t slot = wpieces[w];
    const int64_t base = int64_t(w) * kTileWindow;
    for (int t = 0; t < nt; t++)
      for (const PieceRec& pr : recs[t][w]) {
        if (fill + pr.len > kTilePiece) {
          c++;
          fill = 0;
          chunks[c] = {w, static_cast<int32_t>(slot), 0, 0};
        }
        chunks[c].head |= 1ull << fill;
        for (int j = 0; j < pr.len; j++) {
          src[c * 64 + fill + j] = static_cast<uint16_t>(nbr[pr.begin + j] - base);
          chunks[c].valid |= 1ull << (fill + j);
        }
        fill += pr.len;
        pslot[pnext[pr.row]++] = static_cast<int32_t>(slot);
        slot++;
      }
  }
  recs.clear();
  // 6. Row-kernel items over the lists: heavy rows whole, long rows in
  //    kPrSegment-edge segments; a split row's pieces are spread over its
  //    segments in contiguous ranges.
  std::vector<TItem> items;
  std::vector<PrSuper> supers;
  int32_t slots = 0;
  for (int64_t r = 0; r < n; r++) {
    const int64_t b = loff[r], e = loff[r + 1];
    if (e - b < ctx->heavy_threshold) continue;
    const int64_t pb = poff[r], pe = poff[r + 1];
    if (e - b < kPrSplit) {
      items.push_back({static_cast<int32_t>(r), -1, b, e, pb, pe});
      continue;
    }
    const int64_t k = (e - b + kPrSegment - 1) / kPrSegment;
    PrSuper sp{static_cast<int32_t>(r), slots, static_cast<int32_t>(k), 0};
    for (int64_t j = 0; j < k; j++) {
      const int64_t sb = b + j * kPrSegment, se = std::min(e, sb + kPrSegment);
      items.push_back({static_cast<int32_t>(r), slots++, sb, se, pb + (pe - pb) * j / k, pb + (pe - pb) * (j + 1) / k});
    }
    supers.push_back(sp);
  }
  // 7. Upload.
  int rc;
  std::vector<int64_t> twin(wchunks.begin(), wchunks.end());
  if ((rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_tl_off), loff, "tiled PR row lists")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_tl_nbr), lnbr, "tiled PR row lists")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_tl_poff), poff, "tiled PR piece lists")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_tl_pslot), pslot, "tiled PR piece lists")) ||
      (rc = upload_vec(ctx, &ctx->d_titems, items, "tiled PR items")) ||
      (rc = upload_vec(ctx, &ctx->d_tsupers, supers, "tiled PR split rows")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_tsrc), src, "tiled PR chunk sources")) ||
      (rc = upload_vec(ctx, &ctx->d_tchunks, chunks, "tiled PR chunks")) ||
      (rc = upload_vec(ctx, reinterpret_cast<void**>(&ctx->d_twin), twin, "tiled PR windows")))
    return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_tpartial, 16 * std::max<int64_t>(1, slots), "tiled PR segment partials")) ||
      (rc = dev_alloc(ctx, &ctx->d_piece_part, 8 * std::max<int64_t>(1, npieces), "tiled PR piece partials")))
    return rc;
  ctx->n_titems = static_cast<int64_t>(items.size());
  ctx->n_tsuper = static_cast<int64_t>(supers.size());
  ctx->n_chunks = nchunks;
  ctx->n_windows = nwin;
  ctx->n_pieces = npieces;
  ctx->list_edges = nlist;
  ctx->tile_edges = m - nlist;
  ctx->tiles_ready = true;
  return DMM_OK;
}

}  // namespace dmm
