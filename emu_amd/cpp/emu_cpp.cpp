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
                                     const char* path, const char* dump_path, int quiet, double out[6], char* errbuf,
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
        opt.quiet = quiet != 0;
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
