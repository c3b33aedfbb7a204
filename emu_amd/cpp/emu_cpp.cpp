// emu_cpp.cpp -- the C entry points of emu_cpp.h over the C++ layer.
#include "emu_cpp.h"

#include <cstring>
#include <string>
#include <vector>

#include "Emu.hpp"

namespace {

void put(char* buf, int len, const std::string& s)
{
    if (!buf || len <= 0) return;
    std::strncpy(buf, s.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
}

std::vector<std::string> list(const char* const* v, int n)
{
    std::vector<std::string> out;
    for (int i = 0; i < n; i++)
        if (v && v[i]) out.emplace_back(v[i]);
    return out;
}

} // namespace

extern "C" int emu_cpp_read_params(const char* inputs_file, const char* const* overrides, int noverrides,
                                   emu_params* params, emu_init_config* init, char* errbuf, int errlen)
{
    if (!inputs_file) return EMU_ERR_ARG;
    emu::set_throw_on_error(true);
    try {
        emu::TestParams p;
        p.Initialize(inputs_file, list(overrides, noverrides));
        if (params) *params = p.c_params();
        if (init) *init = p.init_config();
        put(errbuf, errlen, "");
        return EMU_OK;
    } catch (const emu::Error& e) {
        put(errbuf, errlen, e.what());
        return e.code ? e.code : EMU_ERR_ARG;
    } catch (const std::exception& e) {
        put(errbuf, errlen, e.what());
        return EMU_ERR_ARG;
    }
}

extern "C" int emu_cpp_evolve_flavor(const char* inputs_file, const char* const* overrides, int noverrides,
                                     const char* path, const char* dump_path, int flags, double out[6], char* errbuf,
                                     int errlen)
{
    if (!inputs_file) return EMU_ERR_ARG;
    emu::set_throw_on_error(true);
    try {
        emu::TestParams p;
        p.Initialize(inputs_file, list(overrides, noverrides));
        emu::EvolveOptions opt;
        if (path) opt.path = path;
        if (dump_path) opt.dump_path = dump_path;
        opt.quiet = (flags & 1) != 0;
        opt.write_plotfiles = (flags & 2) != 0;
        const emu::EvolveResult r = emu::evolve_flavor(&p, opt);
        if (out) {
            out[0] = r.steps;
            out[1] = r.time;
            out[2] = r.dt;
            out[3] = r.run_time_s;
            out[4] = r.fom;
            out[5] = (double)r.np;
        }
        put(errbuf, errlen, "");
        return EMU_OK;
    } catch (const emu::Error& e) {
        put(errbuf, errlen, e.what());
        return e.code ? e.code : EMU_ERR_ARG;
    } catch (const std::exception& e) {
        put(errbuf, errlen, e.what());
        return EMU_ERR_ARG;
    }
}

extern "C" int emu_cpp_write_plotfile(const char* dir, int nf, const int ncell[3], const double plo[3],
                                      const double phi[3], double time, int step, const double* valid_grid,
                                      int64_t np, const int64_t* ids, const double* pos, const double* rdata,
                                      char* errbuf, int errlen)
{
    if (!dir || !ncell || !plo || !phi || !valid_grid || (nf != 2 && nf != 3) || np < 0 ||
        (np > 0 && (!ids || !pos || !rdata)))
        return EMU_ERR_ARG;
    emu::set_throw_on_error(true);
    try {
        emu::PlotFileSpec sp{};
        sp.nf = nf;
        for (int d = 0; d < 3; d++) {
            sp.ncell[d] = ncell[d];
            sp.plo[d] = plo[d];
            sp.phi[d] = phi[d];
        }
        sp.time = time;
        sp.step = step;
        emu::write_plotfile_grid(dir, sp, valid_grid);
        if (np > 0) emu::write_plotfile_particles(dir, "neutrinos", sp, np, ids, pos, rdata);
        put(errbuf, errlen, "");
        return EMU_OK;
    } catch (const emu::Error& e) {
        put(errbuf, errlen, e.what());
        return e.code ? e.code : EMU_ERR_ARG;
    } catch (const std::exception& e) {
        put(errbuf, errlen, e.what());
        return EMU_ERR_ARG;
    }
}

extern "C" int emu_cpp_read_plotfile(const char* dir, int nf, double* time, int* step, int64_t* np, int64_t capacity,
                                     int64_t* ids, double* pos, double* rdata, char* errbuf, int errlen)
{
    if (!dir || !time || !step || !np || (nf != 2 && nf != 3)) return EMU_ERR_ARG;
    emu::set_throw_on_error(true);
    try {
        emu::read_plotfile_header(dir, *time, *step);
        std::vector<int64_t> vi;
        std::vector<double> vp, vr;
        const int64_t n = emu::read_plotfile_particles(dir, "neutrinos", nf, vi, vp, vr);
        *np = n;
        if (capacity >= n && n > 0) {
            if (!ids || !pos || !rdata) return EMU_ERR_ARG;
            std::memcpy(ids, vi.data(), vi.size() * sizeof(int64_t));
            std::memcpy(pos, vp.data(), vp.size() * sizeof(double));
            std::memcpy(rdata, vr.data(), vr.size() * sizeof(double));
        }
        put(errbuf, errlen, "");
        return EMU_OK;
    } catch (const emu::Error& e) {
        put(errbuf, errlen, e.what());
        return e.code ? e.code : EMU_ERR_ARG;
    } catch (const std::exception& e) {
        put(errbuf, errlen, e.what());
        return EMU_ERR_ARG;
    }
}
