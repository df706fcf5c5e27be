// fv_tables.cpp -- host-side tables of the step and adjoint launches (see
// fv_tables.h).  No HIP call: this file builds alone on any host.
#include "fv_tables.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fvh {

// Cells whose cfl reaches cfl[0] under the reference's fold
// (common/main.cu:503-524): follow each cell's value forward through the
// launches.  A launch of S threads moves slot p to p - S when S <= p < 2S and
// strands slots >= 2S.
std::vector<uint8_t> fold_subset(int nx, int ny) {
  const long N = (long)nx * ny;
  std::vector<long> strides;
  long sizered = N;
  while (sizered > 32) { sizered /= 2; strides.push_back((sizered / 16) * 16); }
  while (sizered > 2) { sizered /= 2; strides.push_back(sizered); }
  // Walk the folds backwards from cfl[0].  A fold of stride S sets slot
  // i < S to max(slot i, slot i + S) and leaves the slots >= S as they are,
  // so a later fold can read such a slot again (when its stride is more than
  // S/2, as where the two loops meet).  Slots past the grid hold no cell.
  std::vector<uint8_t> m(N, 0);
  m[0] = 1;
  for (auto it = strides.rbegin(); it != strides.rend(); ++it) {
    const long S = *it;
    for (long i = 0; i < S && i + S < N; ++i)
      if (m[i]) m[i + S] = 1;
  }
  return m;
}

// Parameters every entry point that sets up an instance accepts.
int check_params(const fivonagi_params& q, bool f64, std::string* why) {
  auto fail = [&](int code, const char* msg) {
    if (why) *why = msg;
    return code;
  };
  if (q.nx < 8 || q.ny < 8) return fail(FIVONAGI_ERR_ARG, "nx, ny must be >= 8");
  // the kernels address a field with 32-bit byte offsets: at most 4 GiB
  // per field (2^30 cells in FP32, 2^29 in FP64)
  if ((long)q.nx * q.ny > (f64 ? (1L << 29) : (1L << 30)))
    return fail(FIVONAGI_ERR_ARG, "grid too large (a field must fit 4 GiB)");
  if (!(q.dx > 0) || !(q.dy > 0) || !(q.dt0 > 0) || !(q.cfl_wish > 0))
    return fail(FIVONAGI_ERR_ARG, "dx, dy, dt0, cfl_wish must be > 0");
  if (q.bc < 0 || q.bc > 3) return fail(FIVONAGI_ERR_ARG, "unknown boundary kind");
  if (q.bc == FIVONAGI_BC_HIFU &&
      !(q.hifu_f > 0 && q.hifu_frec > 0 && q.hifu_a >= 0 && std::isfinite(q.hifu_ampl)))
    return fail(FIVONAGI_ERR_ARG, "hifu: hifu_f, hifu_frec must be > 0, hifu_a >= 0");
  if (q.cfl_mode < 0 || q.cfl_mode > 1) return fail(FIVONAGI_ERR_ARG, "unknown cfl mode");
  if (q.md_on != 0 && q.md_on != 1) return fail(FIVONAGI_ERR_ARG, "md_on must be 0 or 1");
  if (q.md_on && !(q.md_vx >= 0 && q.md_vy >= 0 && std::isfinite(q.md_vx) &&
                   std::isfinite(q.md_vy) && std::isfinite(q.md_t)))
    return fail(FIVONAGI_ERR_ARG, "moving window: md_vx, md_vy must be finite and >= 0");
  return 0;
}

// The host side of one instance: every scalar field of its descriptor (the
// device pointers are left null) and the tables its device copies are made
// from -- edge drive or HIFU aperture, CFL fold subset.
void host_inst(const fivonagi_params& q, bool f64, HostInst& h) {
  fv::InstDesc& d = h.d;
  std::memset(&d, 0, sizeof(d));
  const size_t n = (size_t)q.nx * q.ny;
  d.fp64 = f64 ? 1 : 0;
  d.hbeta_d = q.beta - 1.0;
  d.hifu_k_d = q.hifu_frec / 4.0;
  // edge drive amplitude/phase folded into amp*cos(phase), amp*sin(phase);
  // the HIFU aperture's AMPL and phase delay ISPW*phi_j the same way
  h.ec.assign(q.ny, 0.0);
  h.es.assign(q.ny, 0.0);
  h.tab.clear();
  const bool hifu = q.bc == FIVONAGI_BC_HIFU;
  const double ispw = 2.0 * 3.14159265358979323846 * q.hifu_frec;   // ISPW
  if (hifu) h.tab.assign(4 * (size_t)q.ny, 0.0);
  for (int j = 0; j < q.ny; ++j) {
    if (hifu) {
      // apps/hifu-beam/boundary.h:79-90
      const int width = (int)(q.hifu_a / q.dy);
      const int jp = j - 2;
      if (jp >= 0 && jp < width) {
        const double F = q.hifu_f, jp2 = (double)jp * jp;
        const double nd = std::sqrt(jp2 + F * F);
        const double phi = (q.dx * jp2) / (2.0 * F);
        h.ec[j] = q.hifu_ampl * std::cos(ispw * phi);
        h.es[j] = q.hifu_ampl * std::sin(ispw * phi);
        h.tab[4 * (size_t)j + 0] = phi;
        h.tab[4 * (size_t)j + 1] = F / nd;
        h.tab[4 * (size_t)j + 2] = -jp / nd;
        h.tab[4 * (size_t)j + 3] = 1.0;
      }
      continue;
    }
    const double a = q.edge_amp ? q.edge_amp[j] : 0.0;
    const double ph = q.edge_phase ? q.edge_phase[j] : 0.0;
    h.ec[j] = a * std::cos(ph);
    h.es[j] = a * std::sin(ph);
  }
  d.hifu_k = (float)(q.hifu_frec / 4.0);
  // CFL subset
  d.cfl_mask = fv::kCflAll;
  h.fold_bits.clear();
  if (q.cfl_mode == FIVONAGI_CFL_REFERENCE_FOLD) {
    std::vector<uint8_t> m = fold_subset(q.nx, q.ny);
    bool even = true;
    for (size_t o = 0; o < n && even; ++o) even = (m[o] != 0) == ((o & 1) == 0);
    if (even) {
      d.cfl_mask = fv::kCflEven;
    } else {
      h.fold_bits.assign((n + 31) / 32, 0u);
      for (size_t o = 0; o < n; ++o)
        if (m[o]) h.fold_bits[o >> 5] |= 1u << (o & 31);
      d.cfl_mask = fv::kCflBitmap;
    }
  }
  d.dt_seq = nullptr;
  d.dt_seq_len = 0;
  d.dx = q.dx; d.dy = q.dy; d.delta = q.delta;
  d.cfl_max = q.cfl_max; d.cfl_wish = q.cfl_wish; d.omega = q.omega;
  d.inv_dx = 1.0 / q.dx; d.inv_dy = 1.0 / q.dy;
  d.dif_x = q.delta / (q.dx * q.dx); d.dif_y = q.delta / (q.dy * q.dy);
  d.omega_turns = (hifu ? ispw : q.omega) / 6.283185307179586476925286766559;
  d.tay_rhoa = q.tay_rhoa; d.tay_c = q.tay_c;
  d.tay_cos = std::cos(q.tay_theta); d.tay_sin = std::sin(q.tay_theta);
  d.tay_yoff = q.tay_yoff; d.tay_jc = q.tay_jc;
  d.nx = q.nx; d.ny = q.ny; d.bc = q.bc;
  h.taylor_ring = q.bc == FIVONAGI_BC_TAYLOR ? 4 * q.nx + 4 * (q.ny - 4) : 0;
  d.beta2m2 = (float)(2.0 * (q.beta - 1.0));
  d.md_on = q.md_on;
  d.md_vx = q.md_vx; d.md_vy = q.md_vy; d.md_t = q.md_t;
  h.ext = hifu || q.md_on;
}

// A table of doubles as the bytes of the context's precision.
void table_bytes(const std::vector<double>& v, bool f64, char* out) {
  if (f64) {
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  } else {
    for (size_t k = 0; k < v.size(); ++k) {
      const float f = (float)v[k];
      std::memcpy(out + sizeof(float) * k, &f, sizeof(float));
    }
  }
}

bool cfl_counted(int cfl_mask, const std::vector<uint32_t>& fold_bits, int nx, int i, int j) {
  const size_t o = (size_t)i + (size_t)j * (size_t)nx;
  if (cfl_mask == fv::kCflEven) return (o & 1u) == 0u;
  if (cfl_mask == fv::kCflBitmap) return (fold_bits[o >> 5] >> (o & 31u)) & 1u;
  return true;
}

// Rows per work item: the smallest height (never below 12) whose total wave
// count over all instances fits one resident round of the device
// (CUs x resident step waves per CU): every wave then starts at once, and
// taller work items repeat fewer halo rows.
int auto_rows(const std::vector<fivonagi_params>& ps, long target_waves) {
  for (int H = 12; H <= 4096; ++H) {
    long waves = 0;
    for (const auto& p : ps) {
      const long strips = (p.nx + fv::kStripCols - 1) / fv::kStripCols;
      waves += strips * std::max(1L, std::lround((double)p.ny / H));
    }
    if (waves <= target_waves) return H;
  }
  return 4096;
}

// Row boundaries of one strip cut into `chunks` work items.  With skew > 0
// the items get `levels` height levels (item j: level (j + rot) % levels)
// with weights falling linearly from 1 + skew to 1 - skew (mean 1), so the
// strip keeps its number of items, some taller and some shorter.  rot (the
// strip index) spreads the levels a strip has one more item of evenly over
// the strips, so every level has the same number of items.
std::vector<int> chunk_bounds(int ny, int chunks, float skew, int levels, int rot) {
  const int L = std::max(1, std::min(levels, chunks));
  std::vector<double> cum(chunks + 1, 0.0);
  for (int j = 0; j < chunks; ++j) {
    const int level = (j + rot) % L;
    const double w = (L > 1) ? 1.0 + (double)skew * (L - 1 - 2 * level) / (L - 1) : 1.0;
    cum[j + 1] = cum[j] + std::max(w, 0.05);
  }
  std::vector<int> b(chunks + 1);
  for (int j = 0; j <= chunks; ++j) b[j] = (int)std::lround((double)ny * cum[j] / cum[chunks]);
  b[0] = 0;
  b[chunks] = ny;
  // no empty item; the ends stay at 0 and ny (chunks <= ny)
  for (int j = 1; j < chunks; ++j) b[j] = std::max(b[j], b[j - 1] + 1);
  for (int j = chunks - 1; j >= 1; --j) b[j] = std::min(b[j], b[j + 1] - 1);
  return b;
}

// Launch order of the work items.  Workgroup b of a launch runs on XCD b % 8
// and takes item x*q + min(x, rem) + (b >> 3) (fv_step_kernel), so XCD x
// works through the run of items [x*q + min(x, rem), ...) in order, and within
// a run the hardware starts items in that order, one per SIMD, then the next
// round of SIMD slots.  Each run is sorted by height, tallest first: the
// waves that start first (and, as the oldest on their SIMD, issue first) get
// the most rows, so the SIMDs' waves finish together instead of leaving one
// wave alone at the end; on launches of several rounds the short items come
// last.  Stable, so equal heights keep the strip order (neighbouring items,
// which share halo rows, stay on the same XCD).
void order_for_launch(std::vector<fv::WorkItem>& items) {
  const int n = (int)items.size();
  const int q = n >> 3, rem = n & 7;
  for (int x = 0; x < 8; ++x) {
    const int b = x * q + std::min(x, rem), e = b + q + (x < rem ? 1 : 0);
    std::stable_sort(items.begin() + b, items.begin() + e,
                     [](const fv::WorkItem& u, const fv::WorkItem& v) {
                       return (u.y1 - u.y0) > (v.y1 - v.y0);
                     });
  }
}

constexpr float kAutoSkew = 0.5f;

int item_chunks(const fivonagi_params& p, int auto_h) {
  const int H = p.rows_per_wave > 0 ? p.rows_per_wave : auto_h;
  return std::max(1, (int)std::lround((double)p.ny / H));
}

void build_work(const std::vector<fivonagi_params>& ps, const std::vector<int>& cfl_mask,
                const std::vector<std::vector<uint32_t>>& fold_bits, long cus, int per_cu,
                int auto_h, float skew, bool rotate, WorkList& out) {
  out.items.clear();
  out.rowmask.clear();
  out.n_items.assign(ps.size(), 0);
  const long simds = 4 * cus;
  const int per_simd = std::max(1, per_cu / 4);
  // height levels = waves each SIMD starts in the launch's first round
  long total = 0;
  for (const auto& p : ps)
    total += (long)((p.nx + fv::kStripCols - 1) / fv::kStripCols) * item_chunks(p, auto_h);
  const int levels = (int)std::max(1L, std::min<long>(per_simd, (total + simds - 1) / simds));
  // automatic: skewed only when every SIMD starts its full complement of
  // waves (there the oldest-first issue order staggers their ends; with
  // fewer waves per SIMD they progress together)
  if (skew < 0.0f) skew = (levels >= 4 && levels == per_simd) ? kAutoSkew : 0.0f;
  out.skew_used = skew;
  out.rotate_used = rotate && skew > 0.0f;
  for (int k = 0; k < (int)ps.size(); ++k) {
    const auto& p = ps[k];
    const int chunks = item_chunks(p, auto_h);
    const int strips = (p.nx + fv::kStripCols - 1) / fv::kStripCols;
    // rotate: the level pattern shifts by one per strip (every level gets the
    // same number of items; strips no longer share their cut rows)
    std::vector<std::vector<int>> bounds_rot;
    for (int r = 0; r < (rotate ? levels : 1); ++r)
      bounds_rot.push_back(chunk_bounds(p.ny, chunks, skew, levels, r));
    int n = 0;
    for (int s = 0; s < strips; ++s) {
      const std::vector<int>& bounds = bounds_rot[rotate ? s % levels : 0];
      for (int ch = 0; ch < chunks; ++ch) {
        fv::WorkItem w;
        w.inst = k;
        w.x0 = s * fv::kStripCols;
        w.y0 = bounds[ch];
        w.y1 = bounds[ch + 1];
        w.mrow = (int)out.rowmask.size();
        w.local = n;
        w.bc = p.bc;
        w.pad = 0;
        for (int j = w.y0; j < w.y1; ++j) {
          uint64_t m = 0;
          for (int l = fv::kHalo; l < fv::kLanes - fv::kHalo; ++l) {
            const int i = w.x0 - fv::kHalo + l;
            if (i >= 1 && i < p.nx && cfl_counted(cfl_mask[k], fold_bits[k], p.nx, i, j))
              m |= (uint64_t)1 << l;
          }
          out.rowmask.push_back(m);
        }
        out.items.push_back(w);
        ++n;
      }
    }
    out.n_items[k] = n;
  }
  order_for_launch(out.items);
}

void build_adj_items(const std::vector<fivonagi_params>& ps, long target_waves,
                     int rows_per_wave, std::vector<fv::AdjItem>& items,
                     std::vector<size_t>& first, std::vector<int32_t>& wave_first) {
  constexpr int kCols = 56;   // output columns per strip (a 4-column halo each side)
  // strip starts; the last one never falls on nx-2 or nx-1 (column nx-3's
  // wave also holds nx-2 and nx-1)
  auto strips = [&](int nx) {
    std::vector<int> b;
    for (int x = 0; x < nx; x += kCols) b.push_back(x);
    if (b.size() >= 2 && nx - b.back() <= 2) b.back() = nx - 3;
    b.push_back(nx);
    return b;
  };
  // a row cut never leaves row 2 apart from rows 0, 1, nor row ny-3 apart
  // from ny-2, ny-1 (a ghost copy and its source in one item)
  auto legal_cut = [](int cut, int ny) {
    if (cut == 1 || cut == 2) return 3;
    if (cut == ny - 2 || cut == ny - 1) return ny;
    return cut;
  };
  items.clear();
  wave_first.assign(1, 0);
  first.assign(ps.size() + 1, 0);
  if (rows_per_wave > 0) {
    // fixed heights: rows_per_wave rows per item, one item per wave
    const int H = std::max(8, rows_per_wave);
    for (size_t k = 0; k < ps.size(); ++k) {
      const auto& p = ps[k];
      const std::vector<int> sb = strips(p.nx);
      first[k] = items.size();
      for (size_t s = 0; s + 1 < sb.size(); ++s)
        for (int y = 0; y < p.ny;) {
          const int y1 = std::min(p.ny, legal_cut(std::min(p.ny, y + H), p.ny));
          items.push_back({(int32_t)k, sb[s], sb[s], sb[s + 1], y, y1});
          wave_first.push_back((int32_t)items.size());
          y = y1;
        }
    }
    first[ps.size()] = items.size();
    return;
  }
  // balanced: the strips of every instance, one after another, are cut into
  // waves of about equal cost, an item costing its rows plus kWarm for the
  // rows the march starts with before the first it emits; a wave may end one
  // strip and start the next (two items).  The least cost per wave whose
  // waves fit target_waves is found by bisection.
  constexpr int kWarm = 4, kMinRows = 8;
  auto pack = [&](long T, bool emit) {
    long waves = 0, left = 0;
    for (size_t k = 0; k < ps.size(); ++k) {
      const auto& p = ps[k];
      const std::vector<int> sb = strips(p.nx);
      if (emit) first[k] = items.size();
      for (size_t s = 0; s + 1 < sb.size(); ++s)
        for (int y = 0; y < p.ny;) {
          if (left - kWarm < kMinRows) {       // start a wave
            if (emit && waves > 0) wave_first.push_back((int32_t)items.size());
            ++waves;
            left = T;
          }
          const long take = std::min<long>(p.ny - y, left - kWarm);
          const int y1 = std::min(p.ny, legal_cut(y + (int)take, p.ny));
          if (emit) items.push_back({(int32_t)k, sb[s], sb[s], sb[s + 1], y, y1});
          left -= (y1 - y) + kWarm;
          y = y1;
        }
    }
    if (emit) {
      first[ps.size()] = items.size();
      wave_first.push_back((int32_t)items.size());
    }
    return waves;
  };
  long total = 0;
  for (const auto& p : ps) total += (long)(strips(p.nx).size() - 1) * (p.ny + kWarm);
  long lo = kWarm + kMinRows, hi = std::max(lo, total + kWarm + kMinRows);
  const long target = std::max(1L, target_waves);
  while (lo < hi) {                          // the least T with pack(T) <= target
    const long mid = lo + (hi - lo) / 2;
    if (pack(mid, false) <= target) hi = mid;
    else lo = mid + 1;
  }
  pack(lo, true);
}

}  // namespace fvh
