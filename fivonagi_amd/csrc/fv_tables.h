// fv_tables.h -- the host-side tables of the step and adjoint launches: work
// items and their CFL lane masks, the reference fold's cell subset, the
// per-instance descriptor fields and row tables, parameter checks.  Plain
// C++ with no HIP call, so it builds and runs on any host (the runtime links
// it into libfivonagi.so; tests/sanitize builds it alone with AddressSanitizer
// and UndefinedBehaviorSanitizer).
#ifndef FIVONAGI_FV_TABLES_H
#define FIVONAGI_FV_TABLES_H

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fivonagi.h"
#include "fv_adjoint.h"
#include "fv_types.h"

namespace fvh {

// Cells whose cfl reaches cfl[0] under the reference's fold
// (common/main.cu:503-524), as a 0/1 byte per cell (offset i + j*nx).
std::vector<uint8_t> fold_subset(int nx, int ny);

// 0, or an error code with *why set: the parameters every entry point that
// sets up an instance accepts.
int check_params(const fivonagi_params& q, bool f64, std::string* why);

// The host side of one instance: every scalar field of its descriptor (the
// device pointers are left null) and the tables its device copies are made
// from.
struct HostInst {
  fv::InstDesc d;                    // scalar fields; device pointers null
  std::vector<double> ec, es, tab;   // edge drive (or aperture) and HIFU row tables
  std::vector<uint32_t> fold_bits;   // kCflBitmap: bit i + j*nx = cell counted
  bool ext = false;                  // needs the extended step kernel
  int taylor_ring = 0;               // ghost ring of a Taylor instance (0: none)
};
void host_inst(const fivonagi_params& q, bool f64, HostInst& h);

// A table of doubles as the bytes of the context's precision.
void table_bytes(const std::vector<double>& v, bool f64, char* out);

// Is cell (i, j) one the CFL reduction counts, for an instance's mask kind
// (fv::CflMask) and bitmap?
bool cfl_counted(int cfl_mask, const std::vector<uint32_t>& fold_bits, int nx, int i, int j);

// Rows per work item: the smallest height (never below 12) whose total wave
// count over all instances fits target_waves.
int auto_rows(const std::vector<fivonagi_params>& ps, long target_waves);

// Row boundaries of one strip cut into `chunks` work items (skewed heights,
// `levels` levels, pattern rotated by rot).
std::vector<int> chunk_bounds(int ny, int chunks, float skew, int levels, int rot);

// Launch order of the work items: each XCD's run sorted by height, tallest first.
void order_for_launch(std::vector<fv::WorkItem>& items);

// The work list of the step kernel.
struct WorkList {
  std::vector<fv::WorkItem> items;
  std::vector<uint64_t> rowmask;     // CFL lane mask per item row (fv::WorkItem::mrow)
  std::vector<int> n_items;          // per instance
  float skew_used = 0.0f;
  bool rotate_used = false;
};
// cus, per_cu: the device's CUs and resident step waves per CU; auto_h: the
// rows per item of instances without their own rows_per_wave; skew < 0:
// automatic.
void build_work(const std::vector<fivonagi_params>& ps, const std::vector<int>& cfl_mask,
                const std::vector<std::vector<uint32_t>>& fold_bits, long cus, int per_cu,
                int auto_h, float skew, bool rotate, WorkList& out);

// Work items of the adjoint's reverse march (fv_adjoint_march.hip): strips
// of 56 output columns cut into row ranges, items in instance and strip
// order; first[k] .. first[k+1] are instance k's items, and wave w runs items
// wave_first[w] .. wave_first[w+1] in turn.  rows_per_wave > 0: items of that
// many rows, one per wave; 0: waves of about equal cost, at most target_waves
// of them, a wave taking one or more items (the end of one strip and the
// start of the next).
void build_adj_items(const std::vector<fivonagi_params>& ps, long target_waves,
                     int rows_per_wave, std::vector<fv::AdjItem>& items,
                     std::vector<size_t>& first, std::vector<int32_t>& wave_first);

}  // namespace fvh

#endif  // FIVONAGI_FV_TABLES_H
