// active_brownian: host driver over the libbrownian C-ABI with the reference's
// command line (cpp/active_brownian.cpp:9-60):
//
//   active_brownian <density> <no_of_colloid> [options]
//
// It runs the start-up sequence and niter MPCD steps on the GPU, then prints
// the elapsed seconds as the reference does (:57-58).  Options (defaults are
// the reference's, cpp/active_brownian.cpp:3-7):
//   --L n --niter n --nsub n --theta deg --v0 x --seed n --mass x
//   --sig x --sigma x --fp32 | --fp64 --maxwell --sphere --audit-every n
//   --observables FILE   every audit, one JSON line: the energies and momentum
//                        the reference computes per step (:39-55) with its
//                        energy_colloid (:51), and the colloid MSD, orientation
//                        autocorrelation and temperatures since start-up
//   --trajectory FILE    every audit, "# step n" and per colloid its position,
//                        velocity and angular velocity in point::print's format
//                        (cpp/parameters.hpp:60), the order of the reference's
//                        commented-out prints (:44-46)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "brownian.h"

static void die(bw_ctx *c, int rc, const char *what) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, c ? bw_last_error(c) : "");
    if (c) bw_ctx_destroy(c);
    exit(1);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <density> <no_of_colloid> [--L n] [--niter n] [--nsub n] [--theta deg] "
                        "[--v0 x] [--seed n] [--mass x] [--sig x] [--sigma x] [--fp32|--fp64] [--maxwell] "
                        "[--sphere] [--audit-every n] [--observables FILE] [--trajectory FILE]\n", argv[0]);
        return 2;
    }
    bw_params p;
    bw_default_params(&p);
    p.density = atoi(argv[1]);
    p.n_colloid = atoi(argv[2]);
    int niter = 21000, audit_every = 0;
    bool sigma_set = false;
    const char *obs_path = NULL, *traj_path = NULL;
    for (int i = 3; i < argc; i++) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : "0";
        if (!strcmp(a, "--L")) p.L = atoi(v), i++;
        else if (!strcmp(a, "--niter")) niter = atoi(v), i++;
        else if (!strcmp(a, "--nsub")) p.n_sub = atoi(v), i++;
        else if (!strcmp(a, "--theta")) p.theta_deg = atof(v), i++;
        else if (!strcmp(a, "--v0")) p.v0 = atof(v), i++;
        else if (!strcmp(a, "--seed")) p.seed = atoi(v), i++;
        else if (!strcmp(a, "--mass")) p.mass_colloid = atof(v), i++;
        else if (!strcmp(a, "--sig")) p.sig_colloid = atof(v), i++;
        else if (!strcmp(a, "--sigma")) p.sigma = atof(v), sigma_set = true, i++;
        else if (!strcmp(a, "--audit-every")) audit_every = atoi(v), i++;
        else if (!strcmp(a, "--observables")) obs_path = v, i++;
        else if (!strcmp(a, "--trajectory")) traj_path = v, i++;
        else if (!strcmp(a, "--fp32")) p.precision = BW_FP32;
        else if (!strcmp(a, "--fp64")) p.precision = BW_FP64;
        else if (!strcmp(a, "--maxwell")) p.vel_law = BW_VEL_MAXWELL;
        else if (!strcmp(a, "--sphere")) p.orient_law = BW_ORIENT_SPHERE;
        else {
            fprintf(stderr, "unknown option %s\n", a);
            return 2;
        }
    }
    if (!sigma_set) p.sigma = 0.80 * p.sig_colloid;
    clock_t begin = clock();
    bw_ctx *c = NULL;
    int rc = bw_ctx_create(&p, &c);
    if (rc) die(c, rc, "bw_ctx_create");
    if ((rc = bw_startup(c))) die(c, rc, "startup");
    FILE *obs = NULL, *traj = NULL;
    if ((obs_path || traj_path) && audit_every <= 0) audit_every = 1;
    if (obs_path && !(obs = fopen(obs_path, "w"))) die(c, BW_ERR_ARG, "--observables: cannot open file");
    if (traj_path && !(traj = fopen(traj_path, "w"))) die(c, BW_ERR_ARG, "--trajectory: cannot open file");
    if (obs && (rc = bw_observables_reset(c))) die(c, rc, "observables");
    const int nc = p.n_colloid;
    double *cp = (double *)calloc(3 * (size_t)(nc + 1), sizeof(double));
    double *cv = (double *)calloc(3 * (size_t)(nc + 1), sizeof(double));
    double *cw = (double *)calloc(3 * (size_t)(nc + 1), sizeof(double));
    int chunk = audit_every > 0 ? audit_every : niter;
    for (int done = 0; done < niter; done += chunk) {
        int k = niter - done < chunk ? niter - done : chunk;
        if ((rc = bw_step(c, k))) die(c, rc, "bw_step");
        unsigned flags = 0;
        if ((rc = bw_status(c, &flags))) die(c, rc, "status");
        if (audit_every > 0) {
            double a[11];
            if ((rc = bw_audit(c, 0, a))) die(c, rc, "audit");
            printf("step %d ke_fluid %.12e ke_colloid %.12e potential %.12e momentum %.3e %.3e %.3e\n", done + k,
                   a[3], a[0], a[2], a[4], a[5], a[6]);
            if (obs) {
                double o[6];
                if ((rc = bw_observables(c, 0, o))) die(c, rc, "observables");
                // cpp/active_brownian.cpp:51: energy_colloid = potential + ke + rotational ke
                fprintf(obs,
                        "{\"step\": %d, \"ke_fluid\": %.17g, \"ke_colloid\": %.17g, \"ang_ke_colloid\": %.17g, "
                        "\"potential\": %.17g, \"energy_colloid\": %.17g, \"momentum\": [%.17g, %.17g, %.17g], "
                        "\"msd\": %.17g, \"orient_corr\": %.17g, \"t_fluid\": %.17g, \"t_colloid\": %.17g, "
                        "\"t_rot\": %.17g}\n",
                        done + k, a[3], a[0], a[1], a[2], a[2] + a[0] + a[1], a[4], a[5], a[6], o[1], o[2], o[3], o[4],
                        o[5]);
                fflush(obs);
            }
            if (traj) {
                if ((rc = bw_download_colloid(c, 0, cp, cv, cw, NULL, NULL, NULL))) die(c, rc, "download");
                fprintf(traj, "# step %d\n", done + k);
                for (int i = 0; i < nc; i++) {
                    const double *q[3] = {cp + 3 * i, cv + 3 * i, cw + 3 * i};
                    for (int t = 0; t < 3; t++) fprintf(traj, "%36.32lf %35.32lf %35.32lf\n", q[t][0], q[t][1], q[t][2]);
                }
                fflush(traj);
            }
        }
    }
    free(cp); free(cv); free(cw);
    if (obs) fclose(obs);
    if (traj) fclose(traj);
    bw_sync(c);
    clock_t end = clock();
    printf("%lf\n", (double)(end - begin) / CLOCKS_PER_SEC);
    bw_ctx_destroy(c);
    return 0;
}
