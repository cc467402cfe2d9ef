/*
 * Host-side checks of libabf's C ABI under AddressSanitizer and UndefinedBehaviorSanitizer.
 *
 * Runs without a GPU: every call below is host arithmetic (workspace sizes, launch plans, batch
 * layouts) or an argument check that returns before any device work.  The sanitizers report
 * out-of-bounds accesses, use after free, signed overflow and misaligned loads in that host
 * code; the program itself checks the documented results.  Built by tests/sanitize/Makefile.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "abf.h"

static int failures = 0;

#define CHECK(cond)                                                                                                \
    do {                                                                                                           \
        if (!(cond)) {                                                                                             \
            fprintf(stderr, "FAILED %s:%d: %s (last error: %s)\n", __FILE__, __LINE__, #cond, abf_last_error());   \
            ++failures;                                                                                            \
        }                                                                                                          \
    } while (0)

static abf_grid_job job(int64_t G, int M, int F, int form, int steer)
{
    abf_grid_job j;
    memset(&j, 0, sizeof j);
    j.geom.num_grid = G;
    j.geom.num_mics = M;
    j.num_freqs = F;
    j.steer_type = steer;
    j.form = form;
    return j;
}

int main(void)
{
    /* workspace sizes of the grid beamformer for every form and a range of shapes */
    const int forms[] = {ABF_FORM_PLAIN_F32, ABF_FORM_MFMA_BF16X3, ABF_FORM_MFMA_F32, ABF_FORM_MFMA_F16X2};
    const int mics[] = {1, 5, 31, 32, 33, 64, 255, 256, 512, 1024, 1100};
    for (int fi = 0; fi < 4; ++fi)
        for (int mi = 0; mi < (int)(sizeof mics / sizeof mics[0]); ++mi) {
            abf_grid_job j = job(40401, mics[mi], 513, forms[fi], ABF_STEER_CLASSIC);
            const size_t n = abf_beamform_workspace_size(&j);
            CHECK(n > 0);
            int32_t kern = -1, cols = -1, fb = -1;
            int64_t wgs = -1;
            const int rc = abf_beamform_plan(&j, &kern, &cols, &fb, &wgs);
            CHECK(rc == ABF_OK || rc == ABF_ERR_UNSUPPORTED);
            if (rc == ABF_OK) CHECK(wgs > 0 && cols > 0 && fb > 0);
        }
    /* every kernel request on every microphone count: the streamed kernels (6, 8) plan for 2..8
     * blocks of 32 microphones (7: 2..16) with the hardware sin/cos and the float32 phase, the
     * others fall back */
    for (int kr = 0; kr <= 8; ++kr)
        for (int mi = 0; mi < (int)(sizeof mics / sizeof mics[0]); ++mi)
            for (int fast = 0; fast < 2 && mics[mi] <= 1024; ++fast) {
                abf_grid_job j = job(40401, mics[mi], 513, ABF_FORM_MFMA_BF16X3, ABF_STEER_TRUE_LEVEL);
                j.tile_kernel = kr;
                j.phase_mode = ABF_PHASE_REFERENCE | (fast ? ABF_PHASE_FAST_TRIG : 0);
                int32_t kern = -1, cols = -1, fb = -1;
                int64_t wgs = -1;
                CHECK(abf_beamform_plan(&j, &kern, &cols, &fb, &wgs) == ABF_OK);
                CHECK(wgs > 0 && cols > 0 && fb > 0 && kern >= 1 && kern <= 8);
                const int mp = (mics[mi] + 31) / 32 * 32;
                const int streamed = kr >= 6 && fast && mp >= 64 && mp <= (kr == 7 ? 512 : 256);
                CHECK((kern >= 6) == streamed);
                CHECK(!streamed || kern == kr);
                CHECK(abf_beamform_workspace_size(&j) > 0);
            }
    {
        abf_grid_job j = job(100, 64, 2, ABF_FORM_MFMA_BF16X3, ABF_STEER_CLASSIC);
        j.tile_kernel = 9;
        CHECK(abf_beamform_plan(&j, NULL, NULL, NULL, NULL) != ABF_OK);
    }
    /* the C4 shape: a CSM above 2^31 bytes */
    {
        abf_grid_job j = job(131072, 1024, 256, ABF_FORM_MFMA_BF16X3, ABF_STEER_TRUE_LEVEL);
        CHECK(abf_beamform_workspace_size(&j) > ((size_t)1 << 30));
    }
    /* the streamed kernels take bands of at most 64 frequencies (their wavenumbers sit in the 64
     * lanes of a register); the row-pair kernel takes the band asked for */
    {
        const int kernels[] = {3, 6, 7, 8};
        for (int i = 0; i < 4; ++i) {
            abf_grid_job j = job(40401, 256, 513, ABF_FORM_MFMA_BF16X3, ABF_STEER_TRUE_LEVEL);
            j.phase_mode = ABF_PHASE_REFERENCE | ABF_PHASE_FAST_TRIG;
            j.tile_kernel = kernels[i];
            j.tile_freqs = 128;
            int32_t kern = -1, cols = -1, fb = -1;
            int64_t wgs = -1;
            CHECK(abf_beamform_plan(&j, &kern, &cols, &fb, &wgs) == ABF_OK);
            CHECK(kern == kernels[i]);
            CHECK(fb == (kern >= 6 ? 64 : 128));
        }
    }
    /* unknown steer type / form: size 0 and a message */
    {
        abf_grid_job j = job(100, 8, 2, ABF_FORM_MFMA_BF16X3, 9);
        CHECK(abf_beamform_workspace_size(&j) == 0);
        CHECK(strstr(abf_last_error(), "steer_type") != NULL);
        j = job(100, 8, 2, 77, ABF_STEER_CLASSIC);
        CHECK(abf_beamform_workspace_size(&j) == 0);
    }
    /* batch layouts: offsets ascending, 256-byte aligned, inside the total */
    {
        abf_grid_job js[5];
        const int bm[5] = {32, 64, 128, 256, 512};
        for (int i = 0; i < 5; ++i) js[i] = job(1000 + 777 * i, bm[i], 1 + 13 * i, ABF_FORM_MFMA_BF16X3, 1 + (i & 3));
        js[2].form = ABF_FORM_PLAIN_F32;
        size_t off[5];
        const size_t total = abf_beamform_batch_workspace_size(js, 5, off);
        CHECK(total > 0);
        for (int i = 0; i < 5; ++i) {
            CHECK(off[i] % 256 == 0);
            CHECK(off[i] < total);
            if (i) CHECK(off[i] > off[i - 1]);
        }
        CHECK(abf_beamform_batch(js, 5, NULL, 0, NULL) != ABF_OK); /* null device pointers */
        CHECK(abf_beamform_batch(js, -1, NULL, 0, NULL) == ABF_ERR_INVALID);
    }
    /* jobs with null device pointers are rejected before any launch */
    {
        abf_grid_job j = job(1000, 64, 8, ABF_FORM_MFMA_BF16X3, ABF_STEER_CLASSIC);
        CHECK(abf_beamform(&j, NULL, 0, NULL) != ABF_OK);
        CHECK(abf_beamform_prepare(&j, NULL, 0, NULL) != ABF_OK);
        CHECK(abf_beamform_run(&j, NULL, 0, NULL) != ABF_OK);
    }
    /* the no-allocation entry point: null job, and a bin larger than its static array (rejected
     * before any device work; the pointers are never dereferenced) */
    CHECK(abf_beamform_noalloc(NULL, NULL) == ABF_ERR_INVALID);
    {
        abf_grid_job j = job(10, 4096, 1, ABF_FORM_PLAIN_F32, ABF_STEER_CLASSIC);
        j.csm = (const void *)256;
        j.wavenumber = (const double *)256;
        j.out = (float *)256;
        j.geom.gpos = j.geom.mpos = (const double *)256;
        CHECK(abf_beamform_noalloc(&j, NULL) == ABF_ERR_WORKSPACE);
        CHECK(strstr(abf_last_error(), "scratch") != NULL);
    }
    /* other workspace queries */
    CHECK(abf_beamform_table_workspace_size(256) > 0);
    CHECK(abf_beamform_eig_workspace_size(256, 16) == 0); /* the eigenvalue form needs no workspace */
    CHECK(abf_beamform_f64_workspace_size(64, 65) > 0);
    CHECK(abf_welch_csm_workspace_size(512, 2048, 64) > 0);
    CHECK(abf_welch_csm_workspace_size(64, 128, 799) > 0);
    CHECK(abf_rfft_workspace_size(2048) > 0);
    CHECK(abf_irfft_workspace_size(2048) > 0);
    CHECK(abf_point_source_workspace_size(64, 51200) > 0);
    CHECK(abf_cleansc_workspace_size(64, 1681, 65, 0) > 0);
    CHECK(abf_clean_workspace_size(64, 1681, 65) > 0);
    /* argument checks that return before any device work */
    CHECK(abf_transfer(NULL, NULL, 4, 3, 1.0, 0, 7, 1, NULL, NULL) == ABF_ERR_INVALID);
    CHECK(abf_transfer(NULL, NULL, 4, 3, 1.0, 5, 1, 1, NULL, NULL) == ABF_ERR_INVALID);
    CHECK(abf_dist_mat(NULL, NULL, 10, 0, NULL, NULL) == ABF_ERR_INVALID);
    CHECK(abf_dist_mat(NULL, NULL, 0, 3, NULL, NULL) == ABF_OK);
    CHECK(abf_calc_csm(NULL, NULL, 3, -1, 0, NULL) == ABF_ERR_INVALID);
    CHECK(abf_calc_csm(NULL, NULL, 3, 70000, 0, NULL) != ABF_OK);
    CHECK(abf_welch_csm(NULL, 0, 2, 32768, 100, 0, NULL, 1.0, (void *)8, 1, (void *)8, (size_t)1 << 30, NULL) ==
          ABF_ERR_UNSUPPORTED);
    CHECK(abf_beamform_table(NULL, NULL, NULL, 10, 8, 1.0, 9, 0, 1.0, 0, NULL, NULL, NULL, 0, NULL) != ABF_OK);
    CHECK(abf_beamform_custom(NULL, NULL, 10, 0, 0, 1.0, NULL, NULL, NULL, 0, NULL) != ABF_OK);
    CHECK(abf_beamform_eig(NULL, NULL, 0, NULL, NULL, 10, 8, 1.0, 1, 0, 1.0, 0, NULL, NULL, NULL, 0, NULL) != ABF_OK);
    CHECK(abf_damas(NULL, 0, NULL, NULL, 20000, 1, 10, 1.0, NULL) == ABF_ERR_UNSUPPORTED);
    CHECK(abf_mix(NULL, 1.0, -1, 1, NULL, NULL) != ABF_OK);
    CHECK(abf_white_noise(NULL, -1, 10, 1.0, NULL, NULL) != ABF_OK);
    CHECK(strstr(abf_version(), "gfx950") != NULL);
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("abi host checks passed\n");
    return 0;
}
