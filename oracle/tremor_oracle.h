/* CPU oracle of the exact template-matching path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and the benchmark's CPU-baseline leg as the
 * checker; never linked into the product library. */
#ifndef TREMOR_ORACLE_H
#define TREMOR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int k;
    int64_t *pos;
    float *knn;
} orc_bsf;

float orc_cc_to_ed(float cc, int m);
float orc_ed_to_cc(float ed, int m);
void orc_znormalize(float *v, int len);
int orc_block_order(const float *q, int m, int *order);
void orc_window_stats_one(const float *w, int m, float *mean_out, float *std_out);
void orc_window_stats(const float *raw, int64_t n, int m, int threads, float *mean, float *std);
float orc_dist(const float *q, const float *raw, int64_t start, int m, float mean, float sd, float bound,
               const int *order, int *blocks_visited);
void orc_dist_all(const float *qz, int m, const float *raw, int64_t n, const float *mean, const float *sd,
                  int threads, float *out);
int orc_select_knn(const uint64_t *pos, const float *d, int64_t count, int k, int merge_offset, uint64_t *out_pos,
                   float *out_d);
void orc_bsf_init(orc_bsf *q, int k);
void orc_bsf_free(orc_bsf *q);
void orc_bsf_insert(orc_bsf *q, float data, int64_t position);
void orc_bsf_insert_offset(orc_bsf *q, float data, int64_t position, int merge_offset);
int orc_scan_knn(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                 const int64_t *offsets, int count, int k, int merge_offset, int threads, uint64_t *out_pos,
                 float *out_d, double *seconds);
int64_t orc_scan_threshold(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                           const int64_t *offsets, int count, float rho, int merge_offset, int threads,
                           int64_t cap, int32_t *out_t, uint64_t *out_pos, float *out_d);

void orc_sax_breakpoints(float *bp);
void orc_paa(const float *z, int m, int segments, float *paa);
void orc_sax_words(const float *raw, int64_t n, int m, int segments, int threads, uint8_t *out);
float orc_lower_bound(const float *bp, const float *paa, const uint8_t *word, int m, int segments);
void orc_lower_bounds(const float *bp, const float *paa, const uint8_t *words, int64_t count, int m, int segments,
                      float *out);
int orc_scan_knn_lb(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                    const int64_t *offsets, int count, int k, int merge_offset, int threads, int segments,
                    uint64_t *out_pos, float *out_d, double *seconds);
int64_t orc_scan_threshold_lb(const float *raw, int64_t n, const float *templates, const int32_t *lengths,
                              const int64_t *offsets, int count, float rho, int merge_offset, int threads,
                              int segments, int64_t cap, int32_t *out_t, uint64_t *out_pos, float *out_d);

#ifdef __cplusplus
}
#endif
#endif
