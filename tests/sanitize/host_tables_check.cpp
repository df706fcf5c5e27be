// host_tables_check.cpp -- the host-side launch tables (fivonagi_amd/csrc/
// fv_tables.cpp) checked against what they must satisfy, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile):
//   fold_subset        the cells that reach cfl[0] when the reference's fold
//                      (common/main.cu:503-524) is simulated cell by cell
//   chunk_bounds       0 .. ny, strictly increasing
//   build_work         every cell of every instance in exactly one work item;
//                      each row's CFL lane mask = the item's output lanes with
//                      i >= 1 that the instance's subset counts; local indices
//                      0 .. n-1 per instance; the launch order a permutation
//   build_adj_items    every cell in exactly one reverse-march item; no strip
//                      or row boundary between a ghost copy and its source
//   check_params, host_inst, table_bytes on the boundary kinds
// Prints "ok" and exits 0, or reports the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../fivonagi_amd/csrc/fv_tables.h"

namespace {

int failures = 0;

#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                  \
      std::fprintf(stderr, "\n");                         \
      if (++failures > 10) std::exit(1);                  \
    }                                                     \
  } while (0)

// the reference fold, run on one cell's value: launches of S threads each
// setting slot i < S to max(slot i, slot i + S) (i + S < N), the strides
// halving (multiples of 16 while above 32, then down to 2)
std::vector<uint8_t> fold_by_simulation(int nx, int ny) {
  const long N = (long)nx * ny;
  std::vector<long> strides;
  long sizered = N;
  while (sizered > 32) { sizered /= 2; strides.push_back((sizered / 16) * 16); }
  while (sizered > 2) { sizered /= 2; strides.push_back(sizered); }
  std::vector<uint8_t> m(N, 0);
  std::vector<uint8_t> v(N);
  for (long c = 0; c < N; ++c) {
    std::fill(v.begin(), v.end(), 0);
    v[c] = 1;
    for (long S : strides)
      for (long i = 0; i < S && i + S < N; ++i) v[i] = std::max(v[i], v[i + S]);
    m[c] = v[0];
  }
  return m;
}

void check_fold() {
  const int shapes[][2] = {{8, 8}, {9, 8}, {13, 11}, {17, 16}, {32, 32}, {33, 31},
                           {40, 48}, {48, 40}, {64, 17}, {57, 23}, {8, 129}, {100, 12}};
  for (const auto& s : shapes) {
    const auto a = fvh::fold_subset(s[0], s[1]);
    const auto b = fold_by_simulation(s[0], s[1]);
    CHECK(a == b, "fold_subset %dx%d differs from the simulated fold", s[0], s[1]);
  }
  // larger grids: the subset holds cell 0 and is not empty past it
  for (const auto& s : std::vector<std::pair<int, int>>{{4096, 1024}, {1000, 999}}) {
    const auto a = fvh::fold_subset(s.first, s.second);
    CHECK(a.size() == (size_t)s.first * s.second && a[0] == 1, "fold_subset %dx%d",
          s.first, s.second);
  }
}

void check_chunks() {
  for (int ny : {8, 9, 40, 1024, 1031})
    for (int chunks : {1, 2, 3, 7, 8})
      for (float skew : {0.0f, 0.5f, 0.9f})
        for (int levels : {1, 2, 4})
          for (int rot = 0; rot < 4; ++rot) {
            if (chunks > ny) continue;
            const auto b = fvh::chunk_bounds(ny, chunks, skew, levels, rot);
            bool ok = (int)b.size() == chunks + 1 && b[0] == 0 && b[chunks] == ny;
            for (int j = 0; j < chunks && ok; ++j) ok = b[j + 1] > b[j];
            CHECK(ok, "chunk_bounds ny %d chunks %d skew %g levels %d rot %d", ny, chunks, skew,
                  levels, rot);
          }
}

fivonagi_params base_params(int nx, int ny, int bc) {
  fivonagi_params p;
  std::memset(&p, 0, sizeof(p));
  p.nx = nx; p.ny = ny;
  p.dx = 0.05; p.dy = 0.05;
  p.beta = 3.5; p.delta = 5e-5;
  p.cfl_max = 1.0; p.cfl_wish = 0.9;
  p.dt0 = 0.01; p.T0 = 0.0;
  p.bc = bc;
  p.cfl_mode = FIVONAGI_CFL_REFERENCE_FOLD;
  p.omega = 1.0;
  if (bc == FIVONAGI_BC_HIFU) {
    p.hifu_f = 40.0; p.hifu_frec = 1.0; p.hifu_a = 2.0; p.hifu_ampl = 1e-3;
  }
  return p;
}

void check_work(const std::vector<fivonagi_params>& ps, long cus, int per_cu, int auto_h,
                float skew, bool rotate) {
  std::vector<fvh::HostInst> hs(ps.size());
  std::vector<int> masks(ps.size());
  std::vector<std::vector<uint32_t>> bits(ps.size());
  for (size_t k = 0; k < ps.size(); ++k) {
    fvh::host_inst(ps[k], false, hs[k]);
    masks[k] = hs[k].d.cfl_mask;
    bits[k] = hs[k].fold_bits;
  }
  fvh::WorkList wl;
  fvh::build_work(ps, masks, bits, cus, per_cu, auto_h, skew, rotate, wl);
  std::vector<std::vector<int>> cover(ps.size());
  for (size_t k = 0; k < ps.size(); ++k) cover[k].assign((size_t)ps[k].nx * ps[k].ny, 0);
  std::map<std::pair<int, int>, int> locals;
  size_t rows = 0;
  for (const auto& w : wl.items) {
    CHECK(w.inst >= 0 && w.inst < (int)ps.size(), "item instance %d", w.inst);
    const auto& p = ps[w.inst];
    CHECK(w.y0 >= 0 && w.y1 <= p.ny && w.y1 > w.y0, "item rows %d..%d of %d", w.y0, w.y1, p.ny);
    CHECK(w.x0 >= 0 && w.x0 < p.nx && w.x0 % fv::kStripCols == 0, "item x0 %d", w.x0);
    CHECK(w.bc == p.bc, "item bc");
    locals[{w.inst, w.local}] += 1;
    for (int j = w.y0; j < w.y1; ++j) {
      const uint64_t m = wl.rowmask.at((size_t)w.mrow + (j - w.y0));
      ++rows;
      for (int l = 0; l < fv::kLanes; ++l) {
        const int i = w.x0 - fv::kHalo + l;
        const bool out = l >= fv::kHalo && l < fv::kLanes - fv::kHalo && i < p.nx;
        if (out) cover[w.inst][(size_t)i + (size_t)j * p.nx] += 1;
        const bool want = out && i >= 1 &&
                          fvh::cfl_counted(masks[w.inst], bits[w.inst], p.nx, i, j);
        CHECK((((m >> l) & 1u) != 0) == want, "lane mask inst %d cell (%d, %d)", w.inst, i, j);
      }
    }
  }
  CHECK(rows == wl.rowmask.size(), "row masks: %zu rows, %zu masks", rows, wl.rowmask.size());
  for (size_t k = 0; k < ps.size(); ++k) {
    for (size_t o = 0; o < cover[k].size(); ++o)
      CHECK(cover[k][o] == 1, "inst %zu cell %zu covered %d times", k, o, cover[k][o]);
    for (int l = 0; l < wl.n_items[k]; ++l)
      CHECK((locals[{(int)k, l}]) == 1, "inst %zu local index %d", k, l);
  }
  size_t total = 0;
  for (int n : wl.n_items) total += n;
  CHECK(total == wl.items.size(), "item counts");
}

void check_adj(const std::vector<fivonagi_params>& ps, long target, int rows) {
  std::vector<fv::AdjItem> items;
  std::vector<size_t> first;
  std::vector<int32_t> wave_first;
  fvh::build_adj_items(ps, target, rows, items, first, wave_first);
  CHECK(first.size() == ps.size() + 1 && first.back() == items.size(), "first indices");
  // waves: every item in exactly one, in order, none empty; at most target of
  // them when the heights are automatic
  CHECK(wave_first.size() >= 2 && wave_first.front() == 0 &&
            wave_first.back() == (int32_t)items.size(), "wave ranges");
  for (size_t w = 0; w + 1 < wave_first.size(); ++w)
    CHECK(wave_first[w + 1] > wave_first[w], "empty wave %zu", w);
  if (rows == 0) {
    CHECK((long)wave_first.size() - 1 <= std::max(1L, target), "%zu waves for %ld",
          wave_first.size() - 1, target);
    // balance: no wave's cost (rows + 4 per item) above the mean of the
    // waves' costs plus one item's warm-up, one minimal item and the two rows
    // a ghost-row cut can add
    long total = 0, worst = 0;
    for (size_t w = 0; w + 1 < wave_first.size(); ++w) {
      long cost = 0;
      for (int i = wave_first[w]; i < wave_first[w + 1]; ++i)
        cost += (items[i].y1 - items[i].y0) + 4;
      total += cost;
      worst = std::max(worst, cost);
    }
    const long waves = (long)wave_first.size() - 1;
    CHECK(worst <= (total + waves - 1) / waves + 4 + 8 + 2 + 4, "wave cost %ld, mean %ld", worst,
          total / waves);
  }
  for (size_t k = 0; k < ps.size(); ++k) {
    const auto& p = ps[k];
    std::vector<int> cover((size_t)p.nx * p.ny, 0);
    for (size_t it = first[k]; it < first[k + 1]; ++it) {
      const auto& a = items[it];
      CHECK(a.inst == (int)k, "adjoint item instance");
      CHECK(a.x0 == a.xo0 && a.xo1 > a.xo0 && a.xo1 - a.xo0 <= 56 && a.xo1 <= p.nx,
            "adjoint strip %d..%d", a.xo0, a.xo1);
      CHECK(a.y1 > a.y0 && a.y0 >= 0 && a.y1 <= p.ny, "adjoint rows %d..%d", a.y0, a.y1);
      // a ghost copy and its source in one wave
      CHECK(!(a.xo0 == 1 || a.xo0 == 2 || a.xo0 == p.nx - 2 || a.xo0 == p.nx - 1),
            "strip boundary at %d (nx %d)", a.xo0, p.nx);
      CHECK(!(a.y0 == 1 || a.y0 == 2 || a.y0 == p.ny - 2 || a.y0 == p.ny - 1),
            "row boundary at %d (ny %d)", a.y0, p.ny);
      for (int j = a.y0; j < a.y1; ++j)
        for (int i = a.xo0; i < a.xo1; ++i) cover[(size_t)i + (size_t)j * p.nx] += 1;
    }
    for (size_t o = 0; o < cover.size(); ++o)
      CHECK(cover[o] == 1, "adjoint inst %zu cell %zu covered %d times", k, o, cover[o]);
  }
}

void check_params_and_inst() {
  std::string why;
  fivonagi_params p = base_params(64, 48, FIVONAGI_BC_EDGE_DRIVE);
  CHECK(fvh::check_params(p, false, &why) == 0, "valid parameters refused: %s", why.c_str());
  fivonagi_params q = p;
  q.nx = 7;
  CHECK(fvh::check_params(q, false, &why) == FIVONAGI_ERR_ARG, "nx 7 accepted");
  q = p; q.dx = 0.0;
  CHECK(fvh::check_params(q, false, &why) == FIVONAGI_ERR_ARG, "dx 0 accepted");
  q = p; q.bc = 9;
  CHECK(fvh::check_params(q, false, &why) == FIVONAGI_ERR_ARG, "bc 9 accepted");
  q = p; q.nx = 1 << 16; q.ny = 1 << 15;
  CHECK(fvh::check_params(q, false, &why) == FIVONAGI_ERR_ARG, "2^31 cells accepted");
  q = base_params(64, 48, FIVONAGI_BC_HIFU); q.hifu_f = 0.0;
  CHECK(fvh::check_params(q, false, &why) == FIVONAGI_ERR_ARG, "hifu_f 0 accepted");
  // tables
  std::vector<double> amp(48), ph(48);
  for (int j = 0; j < 48; ++j) { amp[j] = 0.01 * j; ph[j] = 0.1 * j; }
  p.edge_amp = amp.data();
  p.edge_phase = ph.data();
  fvh::HostInst h;
  fvh::host_inst(p, false, h);
  CHECK(h.ec.size() == 48 && h.es.size() == 48 && h.tab.empty(), "edge tables");
  for (int j = 0; j < 48; ++j)
    CHECK(std::fabs(h.ec[j] - amp[j] * std::cos(ph[j])) < 1e-15 &&
              std::fabs(h.es[j] - amp[j] * std::sin(ph[j])) < 1e-15,
          "edge table row %d", j);
  std::vector<char> bytes(4 * 48);
  fvh::table_bytes(h.ec, false, bytes.data());
  float f5;
  std::memcpy(&f5, bytes.data() + 4 * 5, 4);
  CHECK(f5 == (float)h.ec[5], "table_bytes");
  fvh::HostInst ht;
  fvh::host_inst(base_params(40, 30, FIVONAGI_BC_TAYLOR), false, ht);
  CHECK(ht.taylor_ring == 4 * 40 + 4 * 26, "taylor ring %d", ht.taylor_ring);
  fvh::HostInst hh;
  fvh::host_inst(base_params(64, 128, FIVONAGI_BC_HIFU), true, hh);
  CHECK(hh.tab.size() == 4 * 128 && hh.ext && hh.d.fp64 == 1, "hifu tables");
}

}  // namespace

int main() {
  check_fold();
  check_chunks();
  check_params_and_inst();
  std::mt19937 rng(7);
  std::uniform_int_distribution<int> dn(8, 300);
  for (int trial = 0; trial < 24; ++trial) {
    std::vector<fivonagi_params> ps;
    const int ni = 1 + trial % 4;
    for (int k = 0; k < ni; ++k) {
      fivonagi_params p = base_params(dn(rng), dn(rng), (trial + k) % 4);
      if (k % 3 == 1) p.cfl_mode = FIVONAGI_CFL_ALL;
      if (trial % 5 == 0) p.rows_per_wave = 5 + trial;
      ps.push_back(p);
    }
    const long cus = 1 + trial * 11;
    const int per_cu = 4 * (1 + trial % 4);
    const int auto_h = fvh::auto_rows(ps, cus * per_cu);
    check_work(ps, cus, per_cu, auto_h, trial % 3 == 0 ? -1.0f : 0.5f * (trial % 2), trial % 2);
    check_adj(ps, cus * per_cu / 2 + 1, trial % 6 == 0 ? 9 + trial : 0);
  }
  // the widths where the last strip would start on nx-2 or nx-1
  for (int nx : {57, 58, 113, 114, 169, 170})
    check_adj({base_params(nx, 40, FIVONAGI_BC_ZEROGRAD)}, 64, 0);
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
