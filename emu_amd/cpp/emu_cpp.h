/*
 * emu_cpp.h -- C entry points of libemu_cpp.so, the C++ host layer (Emu.hpp),
 * for callers that cannot link C++ (ctypes, other languages).  Errors that
 * the C++ layer reports through emu::Abort come back as a status code and a
 * message instead of ending the process.
 */
#ifndef EMU_CPP_H_
#define EMU_CPP_H_

#include "emu_amd.h"
#include "emu_init.h"

#ifdef __cplusplus
extern "C" {
#endif

/* TestParams::Initialize on an inputs file plus "key=value" overrides,
 * converted to the library's parameter structs.  No GPU work. */
int emu_cpp_read_params(const char* inputs_file, const char* const* overrides, int noverrides, emu_params* params,
                        emu_init_config* init, char* errbuf, int errlen);

/* evolve_flavor (main.cpp:37-227) through the C++ layer.  path: "plain",
 * "fused-fp64" or "fused-fp32"; dump_path (may be NULL): final particles in
 * the EMUDUMP1 format of Emu.hpp; flags: 1 = quiet, 2 = write the plotfiles
 * (plt<step> in the working directory).  out[6] = steps taken, time, last dt,
 * run time (s), particles advanced per microsecond, particle count. */
int emu_cpp_evolve_flavor(const char* inputs_file, const char* const* overrides, int noverrides, const char* path,
                          const char* dump_path, int flags, double out[6], char* errbuf, int errlen);

/* WritePlotFile's format from host arrays (IO.cpp:11-36): `dir` receives the
 * valid cells valid_grid[GIdx::ncomp][nz][ny][nx] as a single-level AMReX
 * plotfile and, if np > 0, the particles (ids[np], pos[3][np],
 * rdata[PIdx::nattribs][np]) as the "neutrinos" particle checkpoint. */
int emu_cpp_write_plotfile(const char* dir, int nf, const int ncell[3], const double plo[3], const double phi[3],
                           double time, int step, const double* valid_grid, int64_t np, const int64_t* ids,
                           const double* pos, const double* rdata, char* errbuf, int errlen);

/* The time, step and "neutrinos" particles of a plotfile into host arrays
 * (RecoverParticles, IO.cpp:38-63).  *np receives the count; the arrays are
 * filled only when capacity >= *np. */
int emu_cpp_read_plotfile(const char* dir, int nf, double* time, int* step, int64_t* np, int64_t capacity,
                          int64_t* ids, double* pos, double* rdata, char* errbuf, int errlen);

#ifdef __cplusplus
}
#endif

#endif
