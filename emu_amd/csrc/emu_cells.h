// emu_cells.h -- the per-function deposit and gather over particles listed by
// the box their stencil covers (emu_cells.hip).
#pragma once

#include "emu_internal.h"

namespace emu {

// Device scratch of one call: the box list (keys, permutation, counts, starts,
// cursors), the gather's combined fields and the scan's temporary storage.
struct BoxScratchSize {
    size_t ints = 0, comb = 0, scan = 0, total = 0;
};
BoxScratchSize box_scratch_size(int64_t np, int64_t nk, int nf, bool gather);

// List the particles by box, then deposit them into the fab (deposit = true;
// the caller zeroes the deposited components first and folds the ghosts after)
// or gather the RHS of each from it.  `scratch` holds box_scratch_size(...)
// bytes.  Asynchronous on the context stream.
int box_list_and_run(emu_ctx* ctx, const PView& P, const Fab& fab, bool deposit, void* scratch, size_t scratch_bytes);

// The context's scratch, grown to `need` bytes (hipMalloc) on first use.
int box_scratch(emu_ctx* ctx, size_t need, void** out);

// The static scratch of the current device: static_scratch_init (at
// emu_create, with the device current) looks up its address once per device;
// static_scratch checks a call's sizes against its capacity and returns it.
int static_scratch_init();
int static_scratch(emu_ctx* ctx, int64_t np, int64_t ncells, void** out, size_t* bytes);

} // namespace emu
