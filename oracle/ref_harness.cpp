// Harness exposing the reference's own hot-path functions, compiled from the
// reference sources where they lie (ZnormSearch.hpp is header-only;
// iSAXPqueue.cpp builds on its own).  TEST INFRASTRUCTURE ONLY: the result,
// oracle/_ref/libtremor_ref.so, is used by tests and by the fixture generator
// under tests/golden/ to pin the oracle; the product never loads it.
#include "distance_computers/ZnormSearch.hpp"
#include "isax/iSAXPqueue.hpp"

#include <cstdint>
#include <cstring>

extern "C" {

float ref_cc_to_ed(float cc, int m) { return cc_to_ed(cc, m); }
float ref_ed_to_cc(float ed, int m) { return ed_to_cc(ed, m); }
void ref_znormalize(float *v, int len) { znormalize_inplace(v, len); }

int ref_block_order(const float *q, int m, int *order)
{
    const std::vector<int> o = query_block_order(q, m);
    if (!o.empty())
        std::memcpy(order, o.data(), sizeof(int) * o.size());
    return static_cast<int>(o.size());
}

float ref_dist(const float *q, const float *raw, int64_t start, int m, float mean, float sd, float bound,
               const int *order, int *blocks_visited)
{
    return l2_dist_znorm_subsequence(q, raw, static_cast<idx_t>(start), m, mean, sd, bound, order, blocks_visited);
}

int ref_select_knn(const uint64_t *pos, const float *d, int64_t count, int k, int merge_offset, uint64_t *out_pos,
                   float *out_d)
{
    std::vector<KnnCandidate> c;
    c.reserve(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; i++)
        c.emplace_back(pos[i], d[i]);
    const std::vector<KnnCandidate> kept = select_knn_with_exclusion(c, k, merge_offset);
    for (size_t i = 0; i < kept.size(); i++) {
        out_pos[i] = kept[i].first;
        out_d[i] = kept[i].second;
    }
    return static_cast<int>(kept.size());
}

void *ref_bsf_init(int k) { return pqueue_bsf_init(k); }
void ref_bsf_insert_offset(void *q, float d, uint64_t pos, int merge_offset)
{
    pqueue_bsf_insert_offset(static_cast<pqueue_bsf *>(q), d, pos, nullptr, merge_offset);
}
void ref_bsf_read(void *q, float *knn, int64_t *pos)
{
    pqueue_bsf *b = static_cast<pqueue_bsf *>(q);
    for (int i = 0; i < b->k; i++) {
        knn[i] = b->knn[i];
        pos[i] = b->position[i];
    }
}
void ref_bsf_destroy(void *q) { pqueue_bsf_destroy(static_cast<pqueue_bsf *>(q)); }
}
