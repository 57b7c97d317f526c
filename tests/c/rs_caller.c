/*
 * rs_caller.c -- a reference-style caller of the RS_* API (the driver loop of
 * simradar.c:1011-1074, reduced): configure, populate, then per pulse set the
 * beam, make the pulse, read the IQ and advance time.  Compiled against
 * include/rs.h and linked with libsimradar_amd.so by tests/test_host.py; with
 * no GPU, RS_init_verbose returns NULL and the program exits with status 3.
 */
#include <stdio.h>

#include "rs.h"

int main(int argc, char **argv) {
    const int pulses = argc > 1 ? 10 : 3;
    RSHandle *H = RS_init_verbose(0);
    if (H == NULL) {
        fprintf(stderr, "rs_caller: RS_init_verbose returned NULL\n");
        return 3;
    }
    RS_set_concept(H, RSSimulationConceptDraggedBackground | RSSimulationConceptBoundedParticleVelocity |
                          RSSimulationConceptUniformDSDScaledRCS);
    RS_set_antenna_params(H, 1.0f, 44.5f);
    RS_set_tx_params(H, 0.2e-6f, 50.0e3f);
    RS_set_prt(H, 1.0e-3f);
    RS_set_lambda(H, 0.1f);
    RS_set_sampling_spacing(H, 15.0f, 1.0f, 1.0f);
    RS_set_dsd_to_mp(H);                          /* simple_ppi.c:37 */
    RS_set_vel_data_to_config(H, LESConfigSuctionVortices);
    RS_set_scan_box(H, RS_suggest_scan_domain(H));
    RS_set_random_seed(H, 19760520);
    RS_set_population(H, 65536);
    RS_populate(H);
    double power = 0.0;
    for (int k = 0; k < pulses; k++) {
        RS_set_beam_pos(H, -1.0f + 0.01f * (float)k, 3.0f);
        RS_make_pulse(H);
        RS_download_pulse_only(H);
        for (uint32_t g = 0; g < H->params.range_count; g++)
            power += (double)H->pulse[g].s0 * H->pulse[g].s0 + (double)H->pulse[g].s1 * H->pulse[g].s1;
        RS_advance_time(H);
    }
    RS_download(H);
    printf("rs_caller: %zu scatterers, %u gates, %d pulses, mean H power %.6e\n", H->num_scats,
           H->params.range_count, pulses, power / (pulses * (double)H->params.range_count));
    RS_free(H);
    return 0;
}
