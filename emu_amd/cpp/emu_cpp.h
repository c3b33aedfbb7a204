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
 * the EMUDUMP1 format of Emu.hpp.  out[6] = steps, time, last dt, run time
 * (s), particles advanced per microsecond, particle count. */
int emu_cpp_evolve_flavor(const char* inputs_file, const char* const* overrides, int noverrides, const char* path,
                          const char* dump_path, int quiet, double out[6], char* errbuf, int errlen);

#ifdef __cplusplus
}
#endif

#endif
