/*
 * ch4g_run — the reference's main program (src/main.c) for the streaming
 * lattice-Boltzmann step, on the device lattice through the drop-in
 * functions of include/ch4g_dropin.h.  Same call order as main.c:7-221,
 * particles included.
 *
 *   ch4g_run [define.h] [param.in]      (defaults: ./define.h ./param.in)
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/ch4g_dropin.h"

int main(int argc, char **argv)
{
    const char *define_h = argc > 1 ? argv[1] : "define.h";
    const char *param_in = argc > 2 ? argv[2] : "param.in";
    ch4g_run_params rp;
    /* assign_parameters, processor_splitting (one GPU), allocate_fields, design_lb */
    if (ch4g_read_run_params(define_h, param_in, &rp)) return 0; /* the reference exits 0 on a bad param.in */
    if (ch4g_dropin_init(&rp)) {
        fprintf(stderr, "Error message -> %s. Exit.\n", ch4g_last_error());
        return -1;
    }
    const int fluid = (rp.cfg.flags & CH4G_LB_FLUID) != 0, thermal = (rp.cfg.flags & CH4G_LB_TEMPERATURE) != 0;
    const int forcing = (rp.cfg.flags & (CH4G_FLUID_FORCING_POISEUILLE | CH4G_FLUID_FORCING_CHANNEL |
                                         CH4G_FLUID_FORCING_HIT | CH4G_TEMPERATURE_BUOYANCY)) != 0;
    pop *p = NULL, *rhs_p = NULL, *g = NULL, *rhs_g = NULL, *old = NULL, *eq = NULL; /* state lives on the device */

    const int lagr = rp.lagrange;
    const int twoway = (rp.pcfg.flags & CH4G_LAGRANGE_TWOWAY_MOMENTUM) != 0;
    if (lagr) allocate_particles();
    initial_conditions(rp.resume);
    initialization_forcing();
    /* the first hydro_fields / build_forcing / hydro_fields of main.c:26-48
       are part of initial_conditions on the device lattice */
    if (lagr) initial_conditions_particles(rp.resume);

    int itime = 0;
    double time_now;
    if (rp.output_at_start) {
        dump_averages();
        if (lagr) {
            output_particles();
            dump_particle_averages();
        }
    }
    for (time_now = rp.time_start + rp.time_dt; time_now <= rp.time_end; time_now += rp.time_dt) {
        itime++;
        if (itime % 10 == 0) fprintf(stderr, "time step %d\n", itime);
        if (fluid) {
            compute_advection(p, rhs_p, rp.cfg.tau_u, eq, 'p');
            add_collision(p, rhs_p, rp.cfg.tau_u, eq, 'p');
        }
        if (thermal) {
            compute_advection(g, rhs_g, rp.cfg.tau_t, eq, 'g');
            add_collision(g, rhs_g, rp.cfg.tau_t, eq, 'g');
        }
        if (forcing || twoway) {
            build_forcing();
            if (twoway && itime > 1) add_particle_feedbacks();
            add_forcing();
        }
        boundary_and_pbc_conditions_for_streaming();
        if (fluid) time_stepping(p, rhs_p, old, old, rp.cfg.tau_u, eq, 'p');
        if (thermal) time_stepping(g, rhs_g, old, old, rp.cfg.tau_t, eq, 'g');
        if (fluid) hydro_fields('p');
        if (thermal) hydro_fields('g');
        dump_averages();
        if (lagr) {
            boundary_conditions_hydro();
            for (int s = 0; s < (int)rp.pcfg.time_dt_lagr_factor; s++) {
                interpolate_vector_at_particles(NULL, 'u');
                if (thermal) interpolate_scalar_at_particles(NULL, 't');
                move_particles();
            }
            output_particles();
            dump_particle_averages();
        }
    }
    measure_time();
    if (rp.output_h5) {
        write_pop_h5(); /* main.c:203-209 */
        if (lagr) write_point_particle_h5();
    }
    ch4g_dropin_finalize();
    fprintf(stdout, "The End\n");
    return 0;
}
