/* A C translation unit against include/lagpm.h, as a maintainer's shim in EUNHA (C) would be:
 * it compiles with a C compiler in strict mode, links against liblagpm.so, takes the address of
 * every entry point (a missing one fails the link) and calls the host-only ones. No GPU is used.
 * Prints: ABI version, table entries of the reference law, kick factors, the box size estimate. */
#include "lagpm.h"

#include <stddef.h>
#include <stdio.h>

/* every function of the header */
static void (*const entry_points[])(void) = {
    (void (*)(void))lagpm_create,
    (void (*)(void))lagpm_destroy,
    (void (*)(void))lagpm_last_error,
    (void (*)(void))lagpm_abi_version,
    (void (*)(void))lagpm_stream,
    (void (*)(void))lagpm_synchronize,
    (void (*)(void))lagpm_box_create,
    (void (*)(void))lagpm_box_destroy,
    (void (*)(void))lagpm_box_params,
    (void (*)(void))lagpm_box_set_cosmology,
    (void (*)(void))lagpm_box_np,
    (void (*)(void))lagpm_box_device_bytes,
    (void (*)(void))lagpm_estimate_box_bytes,
    (void (*)(void))lagpm_box_upload,
    (void (*)(void))lagpm_box_download,
    (void (*)(void))lagpm_box_upload_eunha,
    (void (*)(void))lagpm_box_download_eunha_vel,
    (void (*)(void))lagpm_box_download_eunha,
    (void (*)(void))lagpm_prefof,
    (void (*)(void))lagpm_snapshot_read_header,
    (void (*)(void))lagpm_snapshot_read_dm,
    (void (*)(void))lagpm_snapshot_write,
    (void (*)(void))lagpm_eunha_decode,
    (void (*)(void))lagpm_eunha_encode,
    (void (*)(void))lagpm_pmmain,
    (void (*)(void))lagpm_tsc,
    (void (*)(void))lagpm_tsc_fixed,
    (void (*)(void))lagpm_psolver,
    (void (*)(void))lagpm_green_apply,
    (void (*)(void))lagpm_fda4,
    (void (*)(void))lagpm_set_potential,
    (void (*)(void))lagpm_get_gradient,
    (void (*)(void))lagpm_power_spectrum,
    (void (*)(void))lagpm_power_spectrum_noalloc,
    (void (*)(void))lagpm_force_spline,
    (void (*)(void))lagpm_treecorrection,
    (void (*)(void))lagpm_tree_stats_get,
    (void (*)(void))lagpm_tree_accel_sample,
    (void (*)(void))lagpm_tree_nodes_get,
    (void (*)(void))lagpm_box_set_tuning,
    (void (*)(void))lagpm_box_get_tuning,
    (void (*)(void))lagpm_box_set_group_mode,
    (void (*)(void))lagpm_box_get_group_mode,
    (void (*)(void))lagpm_box_autotune,
    (void (*)(void))lagpm_box_timings,
    (void (*)(void))lagpm_evolfactor,
    (void (*)(void))lagpm_hubble_drag,
    (void (*)(void))lagpm_drift,
    (void (*)(void))lagpm_kdk_step,
    (void (*)(void))lagpm_batch_step,
    (void (*)(void))lagpm_set_step_graphs,
    (void (*)(void))lagpm_set_walk_census,
    (void (*)(void))lagpm_set_batch_pipeline
};

int main(void) {
    static double diff[3 * LAGPM_NSPLINE], slope[3 * LAGPM_NSPLINE];
    lagpm_params p = {0};
    lagpm_evolfact ef;
    p.n = 64;
    p.law = LAGPM_LAW_PMFIT;
    p.precision = LAGPM_PREC_FP32;
    p.deconvolve = 0;
    p.mass = 1.0;
    p.efold = 0.9;
    p.epsilon = 0.1;
    p.rsphere = 4.0;
    p.theta = 0.5;
    p.amax = 48.0;
    p.omep = 0.26;
    p.omeplam = 0.74;
    if (lagpm_abi_version() != LAGPM_ABI_VERSION) return 1;
    if (lagpm_force_spline(&p, diff, slope) != LAGPM_OK) return 2;
    if (lagpm_evolfactor(&p, 1.0, 0.25, &ef) != LAGPM_OK) return 3;
    printf("abi %d entry_points %d\n", lagpm_abi_version(), (int)(sizeof(entry_points) / sizeof(entry_points[0])));
    printf("diff %.17g %.17g %.17g\n", diff[3 * 100], diff[3 * 100 + 1], diff[3 * 100 + 2]);
    printf("slope %.17g\n", slope[3 * 2000]);
    printf("evolfact %.9g %.9g %.9g\n", (double)ef.fact1, (double)ef.fact2, (double)ef.pfact);
    printf("bytes %zu\n", lagpm_estimate_box_bytes(&p, (size_t)64 * 64 * 64));
    /* struct layouts a binding must mirror: size and the offset of the last member */
    printf("layout_params %zu %zu\n", sizeof(lagpm_params), offsetof(lagpm_params, omeplam));
    printf("layout_evolfact %zu %zu\n", sizeof(lagpm_evolfact), offsetof(lagpm_evolfact, fact2_pull));
    printf("layout_tree_stats %zu %zu\n", sizeof(lagpm_tree_stats), offsetof(lagpm_tree_stats, n_pp_within));
    printf("layout_timings %zu %zu\n", sizeof(lagpm_timings), offsetof(lagpm_timings, ms_walk));
    printf("layout_snapshot_info %zu %zu\n", sizeof(lagpm_snapshot_info), offsetof(lagpm_snapshot_info, file_bytes));
    return 0;
}
