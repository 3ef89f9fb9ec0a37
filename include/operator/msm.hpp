// Drop-in for the reference's include/operator/msm.cuh (same names and
// parameter lists; kernels the reference launches with <<<>>> are host
// functions here). Device pointers and layouts are the reference's:
// scalars Fr[length] in Montgomery form; `points` is the pre_comp table (G1:
// 96-byte Jacobian records with Z = 1, level-major; G2: two 96-byte lane
// records per point); `dest` receives one Jacobian point (G1 96 B; G2 192 B in
// libff order, as write_to stores it). Only level 0 of the table is read: the
// result is sum s_i P_i, which is what the reference computes through its
// levels. win_cnt, bucket_cnt and level_stride are accepted and unused.
#pragma once
#include <cstddef>
#include "../lightzk_context.hpp"

// include/operator/msm.cuh:210-228
template <typename devFdT, typename dG1>
void multi_scalar_multiplication_g1(size_t length, devFdT* scalars, dG1* points, dG1* dest, size_t win_siz,
                                    size_t win_cnt, size_t bucket_cnt, size_t level_stride) {
    static_assert(sizeof(devFdT) == 32, "scalars: 32-byte Fr elements");
    (void)win_cnt, (void)bucket_cnt, (void)level_stride;
    lightzk::run([&](hzk_ctx* c) {
        return hzk_msm_g1(c, scalars, points, length, HZK_BASES_REF_LEVELS, win_siz, dest);
    });
}

// include/operator/msm.cuh:230-248
template <typename devFdT, typename dG2>
void multi_scalar_multiplication_g2(size_t length, devFdT* scalars, dG2* points, dG2* dest, size_t win_siz,
                                    size_t win_cnt, size_t bucket_cnt, size_t level_stride) {
    static_assert(sizeof(devFdT) == 32, "scalars: 32-byte Fr elements");
    (void)win_cnt, (void)bucket_cnt, (void)level_stride;
    lightzk::run([&](hzk_ctx* c) {
        return hzk_msm_g2(c, scalars, points, length, HZK_BASES_REF_LEVELS, win_siz, dest);
    });
}

// include/operator/msm.cuh:5-24: dest[i] = k_i * G1 as reference records
// (Z = 1). The reference converts the scalars out of Montgomery form in
// place; these are left unchanged.
template <typename devFdT, typename dG1>
void fix_base_multi_scalar_multiplication_g1(dG1* dest, devFdT* scalars, size_t length) {
    lightzk::run([&](hzk_ctx* c) {
        void* tmp = nullptr;
        int rc = hzk_device_alloc(c, 64 * length, &tmp);
        if (!rc) rc = hzk_fixbase_g1(c, scalars, length, tmp);
        if (!rc) rc = hzk_bases_convert(c, 1, HZK_BASES_AFFINE, HZK_BASES_REF_LEVELS, tmp, dest, length);
        if (!rc) rc = hzk_sync(c);
        if (tmp) hzk_device_free(c, tmp);
        return rc;
    });
}

// include/operator/msm.cuh:26-46 (G2 records, two per point); uni_raw, the
// reference's device copy of the generator, is accepted and unused
template <typename devFdT, typename dG2>
void fix_base_multi_scalar_multiplication_g2(dG2* dest, devFdT* scalars, size_t length, dG2* uni_raw = nullptr) {
    (void)uni_raw;
    lightzk::run([&](hzk_ctx* c) {
        void* tmp = nullptr;
        int rc = hzk_device_alloc(c, 128 * length, &tmp);
        if (!rc) rc = hzk_fixbase_g2(c, scalars, length, tmp);
        if (!rc) rc = hzk_bases_convert(c, 2, HZK_BASES_AFFINE, HZK_BASES_REF_LEVELS, tmp, dest, length);
        if (!rc) rc = hzk_sync(c);
        if (tmp) hzk_device_free(c, tmp);
        return rc;
    });
}

// include/operator/msm.cuh:48-61 / :62-75
template <typename dG1>
void pre_comp_g1(size_t length, dG1* points, size_t gap, size_t level) {
    lightzk::run([&](hzk_ctx* c) { return hzk_precomp(c, 1, points, length, gap, level); });
}
template <typename dG2>
void pre_comp_g2(size_t length, dG2* points, size_t gap, size_t level) {
    lightzk::run([&](hzk_ctx* c) { return hzk_precomp(c, 2, points, length, gap, level); });
}
