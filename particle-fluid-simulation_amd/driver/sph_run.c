/*
 * sph_run: headless C driver of libsph_hip.so, the counterpart of the reference's
 * main/mainLoop/drawFrame (src/main.c:142-199,348-414, src/renderer/renderer.c:383-504)
 * without the window: it fills the reference's default dam-break scene, steps it frame by
 * frame and prints the per-pass timing table the reference prints at exit
 * (main.c:170-198).
 *
 *   sph_run [frames] [particles]
 *
 * Scene: the reference's compile-time constants (main.c:11-30,356-358): box
 * (-140,-40,-40)..(140,40,40), R 2.5, g 100, dt 0.01 with 1 substep, mass 1, pressure
 * multiplier 15000, target density 1, particle radius 0.3, near pressure 0, and the
 * initial block of initializePositionsBuffer (renderer.c:569-608) with velocity
 * (0,0,-100) (renderer.c:642), drawn from libc rand() after srand(0) (main.c:351), both
 * from the library (sph_params_reference_default, sph_scene_fill_reference).
 */
#include <stdio.h>
#include <stdlib.h>

#include "sph_world.h"

static int check(int rc, const char *what) {
    if (rc != SPH_OK) fprintf(stderr, "%s failed (%d): %s\n", what, rc, sph_last_error());
    return rc;
}

int main(int argc, char **argv) {
    const unsigned frames = argc > 1 ? (unsigned)atoi(argv[1]) : 100u;
    const unsigned n = argc > 2 ? (unsigned)atoi(argv[2]) : 300000u;
    sph_params p;
    sph_params_reference_default(&p); /* main.c:11-30,356-358; T = N */

    float *pos = malloc((size_t)n * 16), *vel = malloc((size_t)n * 16);
    if (!pos || !vel) return 1;
    /* initializePositionsBuffer / initializeVelocitiesBuffer after srand(0) */
    if (check(sph_scene_fill_reference(&p, n, 0, SPH_RNG_GLIBC, 0, pos, vel), "sph_scene_fill_reference"))
        return 1;

    sph_world *w = NULL;
    if (check(sph_world_create(&p, n, &w), "sph_world_create")) return 1;
    if (check(sph_world_upload_state(w, pos, vel), "sph_world_upload_state")) return 1;
    if (check(sph_world_step(w, frames), "sph_world_step")) return 1;
    if (check(sph_world_sync(w), "sph_world_sync")) return 1;
    sph_pass_times t;
    if (check(sph_world_pass_times(w, 10, &t), "sph_world_pass_times")) return 1;
    if (check(sph_world_download_state(w, pos, vel, NULL), "sph_world_download_state")) return 1;
    printf("particles %u, frames %u (+10 timed substeps)\n", n, frames);
    printf("hash     %.4f ms\nsort     %.4f ms\nreorder  %.4f ms\ndensity  %.4f ms\n"
           "force    %.4f ms\ntotal    %.4f ms per substep\n",
           t.hash_ms, t.sort_ms, t.reorder_ms, t.density_ms, t.force_ms, t.total_ms);
    printf("particle 0: pos (%.4f, %.4f, %.4f) vel (%.4f, %.4f, %.4f)\n", pos[0], pos[1], pos[2],
           vel[0], vel[1], vel[2]);
    sph_world_destroy(w);
    free(pos);
    free(vel);
    return 0;
}
