/* Prints the size and field offsets of every struct of
 * include/haissinski_mi355x.h as the C compiler lays them out, one
 * "struct field offset" line each (tests/test_abi.py compares them with the
 * ctypes binding). */
#include <stddef.h>
#include <stdio.h>

#include "../include/haissinski_mi355x.h"

#define F(S, f) printf(#S " " #f " %zu\n", offsetof(S, f))
#define Z(S) printf(#S " sizeof %zu\n", sizeof(S))

int main(void) {
    Z(hm_params);
    F(hm_params, E0); F(hm_params, mass); F(hm_params, voltage); F(hm_params, harmonic);
    F(hm_params, radius); F(hm_params, pipe_radius); F(hm_params, alpha_c); F(hm_params, phi_s);
    F(hm_params, freq_rf); F(hm_params, n_turns); F(hm_params, use_wakefield);
    F(hm_params, update_eta); F(hm_params, update_E0); F(hm_params, SR_damping);
    F(hm_params, use_excitation);
    Z(hm_options);
    F(hm_options, normals); F(hm_options, seed); F(hm_options, turn_offset);
    F(hm_options, wake_table); F(hm_options, delta_sigma); F(hm_options, bin_half_width_sigmas);
    F(hm_options, charge); F(hm_options, n_resonators); F(hm_options, timing);
    F(hm_options, resonators); F(hm_options, fft_block); F(hm_options, block_size);
    F(hm_options, monitor); F(hm_options, serial); F(hm_options, fom_stats);
    F(hm_options, tile_pairs); F(hm_options, run_pairs); F(hm_options, bin_whole);
    Z(hm_result);
    F(hm_result, sigma_E); F(hm_result, sigma_z); F(hm_result, E0); F(hm_result, turns_done);
    F(hm_result, drift_skipped); F(hm_result, ms_total); F(hm_result, ms_pass_avg);
    F(hm_result, ms_conv_avg); F(hm_result, pass_launches); F(hm_result, sigma_E0);
    F(hm_result, sigma_z0); F(hm_result, bin_half_width); F(hm_result, mean_z);
    F(hm_result, mean_dE); F(hm_result, corr_z_dE);
    Z(hm_job);
    F(hm_job, params); F(hm_job, opts); F(hm_job, n); F(hm_job, nbins); F(hm_job, particle_seed);
    F(hm_job, sigma_z); F(hm_job, sigma_E); F(hm_job, fail_after_turns); F(hm_job, fail_setup);
    Z(hm_job_result);
    F(hm_job_result, job_id); F(hm_job_result, status); F(hm_job_result, result);
    F(hm_job_result, t_submit); F(hm_job_result, t_admit); F(hm_job_result, t_done);
    F(hm_job_result, centroid_z); F(hm_job_result, centroid_E);
    return 0;
}
