/*
 * emu_init.h -- initial conditions (FlavoredNeutrinoContainer::InitParticles,
 * Source/FlavoredNeutrinoContainerInit.cpp:136-661) written straight into a
 * device SoA particle view.
 */
#ifndef EMU_INIT_H_
#define EMU_INIT_H_

#include "emu_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { EMU_DIRS_SPHERE = 0, EMU_DIRS_CUBE26 = 1, EMU_DIRS_FIBONACCI = 2 };

/* The ParmParse keys of Parameters.H:13-50 that InitParticles reads.  Angles
 * st4_theta/phi are radians as in the decks; st5 holds nnue, nnua, nnux,
 * fxnue, fxnua, fxnux, fynue, fynua, fynux, fznue, fznua, fznux, amplitude. */
typedef struct emu_init_config {
    int simulation_type;      /* 0..5 */
    int nppc[3];
    int direction_set;        /* EMU_DIRS_SPHERE uses uniform_sphere_xyz(nphi_equator) (:9-37) */
    int nphi_equator;
    int ndirections;          /* EMU_DIRS_FIBONACCI */
    int st3_wavelength_fraction_of_domain;
    double st3_amplitude;
    double st4_ndens, st4_theta, st4_phi, st4_fluxfac;
    double st4_ndensbar, st4_thetabar, st4_phibar, st4_fluxfacbar;
    double st4_amplitude;
    double st5[13];
    uint64_t seed;            /* Philox4x32-10 key for every random draw */
} emu_init_config;

/* Number of particles InitParticles creates (cells x nppc x directions). */
int64_t emu_init_count(const emu_params* params, const emu_init_config* cfg);

/* Fill a device SoA view (np = emu_init_count) and, if ids != NULL, device
 * particle ids (1-based, as ParticleType::NextID).  Synchronous. */
int emu_init_particles(emu_ctx* ctx, const emu_init_config* cfg, emu_particles* out, int64_t* ids);

#ifdef __cplusplus
}
#endif

#endif
