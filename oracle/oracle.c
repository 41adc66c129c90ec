/*
 * CPU oracle for the MCL sensor update of uos/tsdf_localization.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker and never as the thing
 * measured or shipped.
 *
 * It restates, in plain C11 + OpenMP, the arithmetic of the reference CPU
 * evaluator and of the map it reads:
 *   - TSDFEvaluator::evaluate / evaluatePose   src/evaluation/tsdf_evaluator.cpp:27-161
 *   - normalisation + pose estimate            src/evaluation/tsdf_evaluator.cpp:159-221
 *   - CudaSubVoxelMap ctor / getIndex / getEntry / setData
 *                                              include/tsdf_localization/cuda/cuda_sub_voxel_map.tcc:5-231
 *   - likelihood value of a TSDF cell          include/tsdf_localization/map/map_util.h:68-73,116-129
 *   - map chunks -> cells and free points      include/tsdf_localization/map/map_util.h:17-154
 *   - SystematicResampler::resample            include/tsdf_localization/resampling/novel_resampling.h:38-73
 *
 * Conventions binding on every function (the reference is built by gcc for
 * baseline x86-64 with -std=c++17 -O3, CMakeLists.txt:5-14):
 *   - float arithmetic left to right, no FMA contraction (this file is compiled
 *     with -ffp-contract=off; baseline x86-64 has no FMA instruction anyway);
 *   - float -> size_t conversions follow what gcc emits on x86-64
 *     (comiss/jae against 2^63, then cvttss2si), written out explicitly in
 *     orc_cvt_f32_u64 so that the out-of-range band the C++ standard leaves
 *     undefined is reproduced instead of being left to the optimiser;
 *   - size_t -> float conversions round to nearest even.
 *
 * Parity status of each function is stated in its comment; every function is
 * pinned by a test under tests/ (closed forms, brute force on tiny inputs,
 * invariants, or the native gcc behaviour it restates).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* x86-64 conversion semantics                                              */
/* ------------------------------------------------------------------------ */

/* cvttss2si with a 64-bit destination: truncation when the value fits in
 * int64, the "integer indefinite" 0x8000000000000000 otherwise (NaN too). */
static inline uint64_t orc_cvtt_f32_i64(float x)
{
    if (x >= -9223372036854775808.0f && x < 9223372036854775808.0f)
        return (uint64_t)(int64_t)x;
    return 0x8000000000000000ull;
}

/* (size_t)x as gcc compiles it for x86-64: values >= 2^63 are shifted down by
 * 2^63 before cvttss2si and the top bit is flipped back; everything else,
 * NaN included, goes straight through cvttss2si.  Used wherever the reference
 * assigns a float to a size_t (cuda_sub_voxel_map.tcc:53-55,58-60,74-76,132-134). */
ORC_EXPORT uint64_t orc_cvt_f32_u64(float x)
{
    if (x >= 9223372036854775808.0f)
        return orc_cvtt_f32_i64(x - 9223372036854775808.0f) ^ 0x8000000000000000ull;
    return orc_cvtt_f32_i64(x);
}

/* The native conversion, compiled by this gcc for this x86-64 host.  Only used
 * by the test that pins orc_cvt_f32_u64 to the compiler's behaviour. */
ORC_EXPORT uint64_t orc_native_cvt_f32_u64(float x)
{
    volatile float v = x;
    return (uint64_t)v;
}

/* size_t -> float, round to nearest even (gcc: cvtsi2ss, with the halving
 * trick for values >= 2^63). */
ORC_EXPORT float orc_cvt_u64_f32(uint64_t v)
{
    return (float)v;
}

/* ------------------------------------------------------------------------ */
/* sin / cos                                                                */
/* ------------------------------------------------------------------------ */
/* The reference calls libm sin/cos on float angles (tsdf_evaluator.cpp:108-113;
 * which overload the unqualified call bound to is not fixed by the sources).
 * The oracle and the device kernels both use this one self-written routine so
 * that the pose matrices are bit-identical on CPU and GPU:
 *   Cody-Waite reduction by pi/2 in double (pi/2 split into 33+33+53 bits, so
 *   k*P1 and k*P2 are exact for |k| < 2^20), then the Taylor series of sin and
 *   cos on |r| <= pi/4 (terms to r^19 / r^20, truncation error < 1e-19), Horner
 *   form, and one rounding to float at the end.
 * Its agreement with glibc sinf/cosf and with (float)sin((double)x) is measured
 * by tests/test_oracle_sincos.py.  Domain: |x| < 1.6e6 rad; outside it, and
 * for NaN/Inf, both results are NaN. */
static const double ORC_TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static const double ORC_PIO2_1 = 0x1.921fb54400000p+0;
static const double ORC_PIO2_2 = 0x1.0b4611a600000p-34;
static const double ORC_PIO2_3 = 0x1.3198a2e037073p-69;

ORC_EXPORT void orc_sincosf(float xf, float* s_out, float* c_out)
{
    double x = (double)xf;
    if (!(fabs(x) < 1.6e6)) {
        *s_out = NAN;
        *c_out = NAN;
        return;
    }
    if (x == 0.0) { /* sin(+-0) = +-0 */
        *s_out = xf;
        *c_out = 1.0f;
        return;
    }
    double kd = rint(x * ORC_TWO_OVER_PI);
    double r = ((x - kd * ORC_PIO2_1) - kd * ORC_PIO2_2) - kd * ORC_PIO2_3;
    double r2 = r * r;
    /* sin r = r + r^3 * (-1/3! + r2*(1/5! + r2*(...))), Horner from the
     * highest order term down */
    double ps = -0x1.2f49b46814157p-57;                  /* -1/19! */
    ps = r2 * ps + 0x1.952c77030ad4ap-49;                /* +1/17! */
    ps = r2 * ps - 0x1.ae7f3e733b81fp-41;                /* -1/15! */
    ps = r2 * ps + 0x1.6124613a86d09p-33;                /* +1/13! */
    ps = r2 * ps - 0x1.ae64567f544e4p-26;                /* -1/11! */
    ps = r2 * ps + 0x1.71de3a556c734p-19;                /* +1/9!  */
    ps = r2 * ps - 0x1.a01a01a01a01ap-13;                /* -1/7!  */
    ps = r2 * ps + 0x1.1111111111111p-7;                 /* +1/5!  */
    ps = r2 * ps - 0x1.5555555555555p-3;                 /* -1/3!  */
    double s = r + (r * r2) * ps;
    double pc = 0x1.e542ba4020225p-62;                   /* +1/20! */
    pc = r2 * pc - 0x1.6827863b97d97p-53;                /* -1/18! */
    pc = r2 * pc + 0x1.ae7f3e733b81fp-45;                /* +1/16! */
    pc = r2 * pc - 0x1.93974a8c07c9dp-37;                /* -1/14! */
    pc = r2 * pc + 0x1.1eed8eff8d898p-29;                /* +1/12! */
    pc = r2 * pc - 0x1.27e4fb7789f5cp-22;                /* -1/10! */
    pc = r2 * pc + 0x1.a01a01a01a01ap-16;                /* +1/8!  */
    pc = r2 * pc - 0x1.6c16c16c16c17p-10;                /* -1/6!  */
    pc = r2 * pc + 0x1.5555555555555p-5;                 /* +1/4!  */
    pc = r2 * pc - 0x1.0000000000000p-1;                 /* -1/2!  */
    double c = 1.0 + r2 * pc;
    int64_t k = (int64_t)kd;
    switch ((int)(k & 3)) {
    case 0: *s_out = (float)s;  *c_out = (float)c;  break;
    case 1: *s_out = (float)c;  *c_out = (float)-s; break;
    case 2: *s_out = (float)-s; *c_out = (float)-c; break;
    default: *s_out = (float)-c; *c_out = (float)s; break;
    }
}

/* ------------------------------------------------------------------------ */
/* Map: CudaSubVoxelMap<float,float> state                                  */
/* ------------------------------------------------------------------------ */
/* Same field order and types as CudaSubVoxelMap::MapCoef
 * (cuda_sub_voxel_map.h:22-50) with IndexT = DataT = float. */
typedef struct orc_coef {
    uint64_t dim_x, dim_y, dim_z;
    float min_x, min_y, min_z;
    float max_x, max_y, max_z;
    float resolution;
    float init_value;
    uint64_t up_dim_x, up_dim_y, up_dim_z, up_dim_2;
    uint64_t sub_dim, sub_dim_2;
    uint64_t grid_occ_size, data_size;
} orc_coef;

ORC_EXPORT size_t orc_coef_size(void) { return sizeof(orc_coef); }

/* sub_voxel_size, cuda_sub_voxel_map.h:15 */
static const float ORC_SUB_VOXEL_SIZE = 1.0f;

/* Constructor arithmetic, cuda_sub_voxel_map.tcc:16-34: float subtraction,
 * fabsf, float division, ceilf, conversion to size_t. */
ORC_EXPORT void orc_coef_init(float min_x, float min_y, float min_z,
                              float max_x, float max_y, float max_z,
                              float resolution, float init_value, orc_coef* c)
{
    c->dim_x = orc_cvt_f32_u64(ceilf(fabsf(max_x - min_x) / resolution));
    c->dim_y = orc_cvt_f32_u64(ceilf(fabsf(max_y - min_y) / resolution));
    c->dim_z = orc_cvt_f32_u64(ceilf(fabsf(max_z - min_z) / resolution));
    c->min_x = min_x; c->min_y = min_y; c->min_z = min_z;
    c->max_x = max_x; c->max_y = max_y; c->max_z = max_z;
    c->resolution = resolution;
    c->init_value = init_value;
    c->up_dim_x = orc_cvt_f32_u64(ceilf(fabsf(max_x - min_x) / ORC_SUB_VOXEL_SIZE));
    c->up_dim_y = orc_cvt_f32_u64(ceilf(fabsf(max_y - min_y) / ORC_SUB_VOXEL_SIZE));
    c->up_dim_z = orc_cvt_f32_u64(ceilf(fabsf(max_z - min_z) / ORC_SUB_VOXEL_SIZE));
    c->up_dim_2 = c->up_dim_x * c->up_dim_y;
    c->sub_dim = orc_cvt_f32_u64(ceilf(ORC_SUB_VOXEL_SIZE / resolution));
    c->sub_dim_2 = c->sub_dim * c->sub_dim;
    c->grid_occ_size = c->up_dim_x * c->up_dim_y * c->up_dim_z;
    c->data_size = 0;
}

/* A map as the oracle reads it: the reference's int32 block table with float
 * voxel values, or the 64-bit extension past the reference's OCC_T = int
 * (cuda_sub_voxel_map.h:13): an int64 block table with the voxels given as
 * int16 TSDF millimetres, read through the likelihood value of each
 * millimetre (map_util.h:116-129; INT16_MIN = never set, |mm| >= t = outside
 * the truncation band: both read init_value). */
typedef struct orc_map {
    const orc_coef* c;
    const int32_t* grid_occ;    /* reference form (NULL: 64-bit form) */
    const float* data;
    const int64_t* grid_occ64;  /* 64-bit form */
    const int16_t* tsdf_mm;
    const float* lut;           /* likelihood of mm in (-t, t) at mm + t - 1 */
    int truncation_mm;
} orc_map;

static uint64_t orc_index_m(const orc_map* m, float x, float y, float z);

/* CudaSubVoxelMap::getIndex, cuda_sub_voxel_map.tcc:51-137 (track_reuse off).
 * Returns data_size for every rejected lookup, like the reference. */
ORC_EXPORT uint64_t orc_get_index(const orc_coef* c, const int32_t* grid_occ,
                                  float x, float y, float z)
{
    orc_map m = {c, grid_occ, NULL, NULL, NULL, NULL, 0};
    return orc_index_m(&m, x, y, z);
}

/* getIndex on a 64-bit block table. */
ORC_EXPORT uint64_t orc_get_index64(const orc_coef* c, const int64_t* grid_occ64, float x, float y, float z)
{
    orc_map m = {c, NULL, NULL, grid_occ64, NULL, NULL, 0};
    return orc_index_m(&m, x, y, z);
}

static uint64_t orc_index_m(const orc_map* m, float x, float y, float z)
{
    const orc_coef* c = m->c;
    /* :53-62 integer-metre bounds check */
    uint64_t grid_x = orc_cvt_f32_u64(x - c->min_x);
    uint64_t grid_y = orc_cvt_f32_u64(y - c->min_y);
    uint64_t grid_z = orc_cvt_f32_u64(z - c->min_z);
    grid_x = orc_cvt_f32_u64(orc_cvt_u64_f32(grid_x) / c->resolution);
    grid_y = orc_cvt_f32_u64(orc_cvt_u64_f32(grid_y) / c->resolution);
    grid_z = orc_cvt_f32_u64(orc_cvt_u64_f32(grid_z) / c->resolution);
    if (grid_x >= c->dim_x || grid_y >= c->dim_y || grid_z >= c->dim_z)
        return c->data_size;
    /* :67-84 block coordinates and position inside the block */
    float x_offset = x - c->min_x;
    float y_offset = y - c->min_y;
    float z_offset = z - c->min_z;
    uint64_t up_x = orc_cvt_f32_u64(x_offset / ORC_SUB_VOXEL_SIZE);
    uint64_t up_y = orc_cvt_f32_u64(y_offset / ORC_SUB_VOXEL_SIZE);
    uint64_t up_z = orc_cvt_f32_u64(z_offset / ORC_SUB_VOXEL_SIZE);
    float sub_pos_x = x_offset - orc_cvt_u64_f32(up_x) * ORC_SUB_VOXEL_SIZE;
    float sub_pos_y = y_offset - orc_cvt_u64_f32(up_y) * ORC_SUB_VOXEL_SIZE;
    float sub_pos_z = z_offset - orc_cvt_u64_f32(up_z) * ORC_SUB_VOXEL_SIZE;
    /* :86-110 block table */
    uint64_t up_index = up_x + up_y * c->up_dim_x + up_z * c->up_dim_2;
    if (up_index >= c->grid_occ_size)
        return c->data_size;
    int64_t sub_index = m->grid_occ ? (int64_t)m->grid_occ[up_index] : m->grid_occ64[up_index];
    if (sub_index < 0)
        return c->data_size;
    /* :130-136 voxel inside the block */
    uint64_t sub_x = orc_cvt_f32_u64(sub_pos_x / c->resolution);
    uint64_t sub_y = orc_cvt_f32_u64(sub_pos_y / c->resolution);
    uint64_t sub_z = orc_cvt_f32_u64(sub_pos_z / c->resolution);
    return (uint64_t)sub_index + sub_x + sub_y * c->sub_dim + sub_z * c->sub_dim_2;
}

/* CudaSubVoxelMap::getEntry, cuda_sub_voxel_map.tcc:140-158. */
static float orc_entry_m(const orc_map* m, float x, float y, float z)
{
    const orc_coef* c = m->c;
    if (m->grid_occ ? m->data == NULL : m->tsdf_mm == NULL)
        return c->init_value;
    uint64_t index = orc_index_m(m, x, y, z);
    if (index >= c->data_size)
        return c->init_value;
    if (m->grid_occ)
        return m->data[index];
    int mm = m->tsdf_mm[index];
    int t = m->truncation_mm;
    if (mm == INT16_MIN || mm <= -t || mm >= t)
        return c->init_value;
    return m->lut[mm + t - 1];
}

ORC_EXPORT float orc_get_entry(const orc_coef* c, const int32_t* grid_occ, const float* data,
                               float x, float y, float z)
{
    if (data == NULL || grid_occ == NULL)
        return c->init_value;
    orc_map m = {c, grid_occ, data, NULL, NULL, NULL, 0};
    return orc_entry_m(&m, x, y, z);
}

/* CudaSubVoxelMap::setData, cuda_sub_voxel_map.tcc:171-231: mark the blocks
 * of every cell, give each marked block an offset in linear block order,
 * fill with init_value, then place every cell through getIndex.
 * grid_occ must hold grid_occ_size entries; on return c->data_size is set and
 * *data_out is a malloc'ed array (caller frees with orc_free).
 * Returns 0, or -1 for "Upper voxel index overflow!". */
ORC_EXPORT int orc_set_data(orc_coef* c, int32_t* grid_occ, const float* cell_xyz,
                            const float* cell_value, size_t num_cells, float** data_out)
{
    *data_out = NULL;
    for (uint64_t i = 0; i < c->grid_occ_size; ++i)
        grid_occ[i] = -1;
    if (num_cells == 0)
        return 0;
    for (size_t i = 0; i < num_cells; ++i) {
        float x_offset = cell_xyz[3 * i + 0] - c->min_x;
        float y_offset = cell_xyz[3 * i + 1] - c->min_y;
        float z_offset = cell_xyz[3 * i + 2] - c->min_z;
        uint64_t up_x = orc_cvt_f32_u64(x_offset / ORC_SUB_VOXEL_SIZE);
        uint64_t up_y = orc_cvt_f32_u64(y_offset / ORC_SUB_VOXEL_SIZE);
        uint64_t up_z = orc_cvt_f32_u64(z_offset / ORC_SUB_VOXEL_SIZE);
        uint64_t up_index = up_x + up_y * c->up_dim_x + up_z * c->up_dim_x * c->up_dim_y;
        if (up_index >= c->grid_occ_size)
            return -1;
        grid_occ[up_index] = 0;
    }
    /* :202-211 keep the running offset in an int: a map whose voxels do not
     * fit in int32 overflows it (undefined); refused here instead (-2) */
    uint64_t sub_size = c->sub_dim * c->sub_dim * c->sub_dim;
    uint64_t blocks = 0;
    for (uint64_t i = 0; i < c->grid_occ_size; ++i)
        blocks += grid_occ[i] >= 0;
    if (sub_size != 0 && blocks > (uint64_t)INT32_MAX / sub_size)
        return -2;
    int64_t current_offset = 0;
    for (uint64_t i = 0; i < c->grid_occ_size; ++i) {
        if (grid_occ[i] >= 0) {
            grid_occ[i] = (int32_t)current_offset;
            current_offset += (int64_t)sub_size;
        }
    }
    c->data_size = (uint64_t)current_offset;
    float* data = (float*)malloc(sizeof(float) * (c->data_size ? c->data_size : 1));
    for (uint64_t i = 0; i < c->data_size; ++i)
        data[i] = c->init_value;
    /* setEntry (:160-169) writes through getIndex without a bounds check; an
     * index outside the array would be a write out of bounds in the
     * reference, the oracle skips it instead. */
    for (size_t i = 0; i < num_cells; ++i) {
        uint64_t idx = orc_get_index(c, grid_occ, cell_xyz[3 * i + 0], cell_xyz[3 * i + 1],
                                     cell_xyz[3 * i + 2]);
        if (idx < c->data_size)
            data[idx] = cell_value[i];
    }
    *data_out = data;
    return 0;
}

ORC_EXPORT void orc_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Likelihood value of a TSDF cell (map_util.h:68-73,116-129)               */
/* ------------------------------------------------------------------------ */
/* init = (expf(-(10*10)/sq/2) / sqrtf(2*sq*pi))^3 with sq = sigma*sigma in
 * float; the exponent argument is computed in double and converted to float
 * at the expf call, the cube is taken in double and stored as float. */
ORC_EXPORT float orc_likelihood_init(float sigma)
{
    float sigma_quad = sigma * sigma;
    float e = expf((float)(-(10.0 * 10.0) / (double)sigma_quad / 2));
    float d = sqrtf((float)((double)(2 * sigma_quad) * M_PI));
    float init = e / d;
    init = init * init * init;
    return init;
}

/* value of a valid cell (weight != 0, |tsdf_mm| < truncation), :124-126 */
ORC_EXPORT float orc_likelihood_value(int16_t tsdf_mm, float sigma)
{
    float sigma_quad = sigma * sigma;
    float tsdf_value = (float)tsdf_mm;
    double value = (double)tsdf_value * 0.001;
    float e = expf((float)(-(value * value) / (double)sigma_quad / 2));
    float d = sqrtf((float)((double)(2 * sigma_quad) * M_PI));
    value = (double)(e / d);
    value = value * value * value;
    return (float)value;
}

/* ------------------------------------------------------------------------ */
/* Map chunks -> cells and free points (map_util.h:17-154)                  */
/* ------------------------------------------------------------------------ */
/* createTSDFMap after the file read: chunk k sits at chunk coordinates
 * pos[3k..3k+2] and holds 64^3 raw TSDFValue words w[v] at flat index
 * v = 4096 i + 64 j + l (voxel x = 64 cx + i, y = 64 cy + j, z = 64 cz + l).
 * A word is {int16 value (low half), int16 weight (high half)} (tsdf.h:11-45).
 * Bounds (:23-64): per axis the min / max of 0 and every chunk coordinate, as
 * floats, then min * 64 * 64 * 0.001 and (max * 64 + 64) * 64 * 0.001 (float
 * products, the last one in double).  Cell position (:114-116):
 * (float)x * 64 (float) * 0.001 (double), stored as float.  Cells keep the
 * chunk order given and x, y, z nesting; the outputs must hold 64^3 entries
 * per chunk.  Returns the counts through *nc / *nf.
 * Parity: pinned by hand-computed bounds and positions, a crafted chunk for
 * the value / weight split and the truncation edge, and a single-voxel chunk
 * for the x-major layout (tests/test_map_chunks.py). */
ORC_EXPORT void orc_chunks_to_cells(const int32_t* pos, const uint32_t* words, size_t num_chunks, float sigma,
                                    orc_coef* c, float* cell_xyz, float* cell_value, size_t* nc,
                                    float* free_xyz, size_t* nf)
{
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    for (size_t k = 0; k < num_chunks; ++k)
        for (int a = 0; a < 3; ++a) {
            float v = (float)pos[3 * k + a];
            lo[a] = v < lo[a] ? v : lo[a];
            hi[a] = v > hi[a] ? v : hi[a];
        }
    for (int a = 0; a < 3; ++a) {
        float l = lo[a] * 64.0f;
        l = l * 64.0f;
        lo[a] = (float)((double)l * 0.001);
        float h = hi[a] * 64.0f + 64.0f;
        h = h * 64.0f;
        hi[a] = (float)((double)h * 0.001);
    }
    orc_coef_init(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], (float)(64 * 0.001), orc_likelihood_init(sigma), c);
    size_t n_cell = 0, n_free = 0;
    for (size_t k = 0; k < num_chunks; ++k) {
        for (size_t v = 0; v < 64 * 64 * 64; ++v) {
            uint32_t w = words[k * 64 * 64 * 64 + v];
            int16_t value = (int16_t)(w & 0xffffu), weight = (int16_t)(w >> 16);
            if (weight == 0)
                continue;
            int xyz[3] = {64 * pos[3 * k] + (int)(v >> 12), 64 * pos[3 * k + 1] + (int)((v >> 6) & 63),
                          64 * pos[3 * k + 2] + (int)(v & 63)};
            float* dst;
            if (fabsf((float)value) < 600.0f) {
                dst = cell_xyz + 3 * n_cell;
                cell_value[n_cell++] = orc_likelihood_value(value, sigma);
            } else {
                dst = free_xyz + 3 * n_free++;
            }
            for (int a = 0; a < 3; ++a)
                dst[a] = (float)((double)((float)xyz[a] * 64.0f) * 0.001);
        }
    }
    *nc = n_cell;
    *nf = n_free;
}

/* ------------------------------------------------------------------------ */
/* Pose matrix of one particle (tsdf_evaluator.cpp:102-145)                 */
/* ------------------------------------------------------------------------ */
/* The reference calls unqualified sin/cos on float angles
 * (tsdf_evaluator.cpp:108-113, 206-213; particle_cloud.cpp:262-267), i.e.
 * std::sin(float) = the C library's sinf.
 * mode 0 and 1: this process's C library sinf/cosf (glibc chooses its FMA
 * or non-FMA build when it loads; the library's TSL_TRIG_GLIBC_FMA and
 * TSL_TRIG_GLIBC_NOFMA restate the two);
 * mode 2: the double-evaluated routine above (the library's TSL_TRIG_DOUBLE). */
static void orc_trig(float a, int mode, float* s, float* c)
{
    if (mode == 2) {
        orc_sincosf(a, s, c);
    } else {
        volatile float av = a;  /* a call into the C library, never a folded constant */
        *s = sinf(av);
        *c = cosf(av);
    }
}

/* L = [R(roll,pitch,yaw) | t] * T with T the row-major scanner->base matrix;
 * writes the 12 used entries of tf_laser. */
ORC_EXPORT void orc_pose_matrix(const float pose[6], const float tf[16], int trig_mode, float L[12])
{
    float sin_alpha, cos_alpha, sin_beta, cos_beta, sin_gamma, cos_gamma;
    orc_trig(pose[3], trig_mode, &sin_alpha, &cos_alpha);
    orc_trig(pose[4], trig_mode, &sin_beta, &cos_beta);
    orc_trig(pose[5], trig_mode, &sin_gamma, &cos_gamma);
    float p[12];
    p[0] = cos_beta * cos_gamma;
    p[4] = cos_beta * sin_gamma;
    p[8] = -sin_beta;
    p[3] = pose[0];
    p[1] = sin_alpha * sin_beta * cos_gamma - cos_alpha * sin_gamma;
    p[5] = sin_alpha * sin_beta * sin_gamma + cos_alpha * cos_gamma;
    p[9] = sin_alpha * cos_beta;
    p[7] = pose[1];
    p[2] = cos_alpha * sin_beta * cos_gamma + sin_alpha * sin_gamma;
    p[6] = cos_alpha * sin_beta * sin_gamma - sin_alpha * cos_gamma;
    p[10] = cos_alpha * cos_beta;
    p[11] = pose[2];
    for (int r = 0; r < 3; ++r) {
        L[4 * r + 0] = p[4 * r] * tf[0] + p[4 * r + 1] * tf[4] + p[4 * r + 2] * tf[8];
        L[4 * r + 1] = p[4 * r] * tf[1] + p[4 * r + 1] * tf[5] + p[4 * r + 2] * tf[9];
        L[4 * r + 2] = p[4 * r] * tf[2] + p[4 * r + 1] * tf[6] + p[4 * r + 2] * tf[10];
        L[4 * r + 3] = p[4 * r] * tf[3] + p[4 * r + 1] * tf[7] + p[4 * r + 2] * tf[11] + p[4 * r + 3];
    }
}

/* ------------------------------------------------------------------------ */
/* Sensor update (tsdf_evaluator.cpp:27-76, 78-157)                         */
/* ------------------------------------------------------------------------ */
typedef struct orc_particle { float first[6]; float second; } orc_particle;  /* particle.h:26-55 */
typedef struct orc_point { float x, y, z; } orc_point;                       /* cuda_evaluator.h:22-39 */

typedef struct orc_params {
    float a_hit, a_range, a_max, max_range;
} orc_params;

/* tsdf_evaluator.h:72-76: inv_max_range_ = 1.0/max_range (double, stored as
 * float), max_range_squared_ = max_range*max_range (float). */
static void orc_derived(const orc_params* p, float* inv_mr, float* mr2)
{
    *inv_mr = (float)(1.0 / (double)p->max_range);
    *mr2 = p->max_range * p->max_range;
}

/* evaluatePose (:27-76) for one pose matrix.  Returns the reference's float
 * sequential sum and the float64 sum of the same float terms. */
static void orc_eval_pose(const orc_map* m, const float L[12], const orc_point* pts, size_t n,
                          const orc_params* prm, float inv_mr, float mr2, float* w32, double* w64)
{
    float eval_sum = 0.0f;
    double eval_sum_d = 0.0;
    for (size_t j = 0; j < n; ++j) {
        float x = pts[j].x, y = pts[j].y, z = pts[j].z;
        float tx = L[0] * x + L[1] * y + L[2] * z + L[3];
        float ty = L[4] * x + L[5] * y + L[6] * z + L[7];
        float tz = L[8] * x + L[9] * y + L[10] * z + L[11];
        float value = orc_entry_m(m, tx, ty, tz);
        float square_dist = x * x + y * y + z * z;
        if (square_dist < mr2)
            value = prm->a_hit * value + prm->a_range * inv_mr;
        else
            value = prm->a_hit * value + prm->a_max;
        eval_sum += value;
        eval_sum_d += (double)value;
    }
    *w32 = eval_sum;
    *w64 = eval_sum_d;
}

/* TSDFEvaluator::evaluate, particle loop only (:85-157).
 * w32[i]: the reference's float weight; w64[i]: float64 sum of the same terms.
 * threads <= 0 uses all cores.  faithful_copy != 0 copies the point vector per
 * particle, as the by-value parameter of evaluatePose does (tsdf_evaluator.h:59). */
static void orc_evaluate_m(const orc_map* m, const orc_particle* particles, size_t num_particles,
                           const orc_point* points, size_t num_points, const float tf[16],
                           const orc_params* prm, int trig_mode, int threads, int faithful_copy,
                           float* w32, double* w64)
{
    float inv_mr, mr2;
    orc_derived(prm, &inv_mr, &mr2);
    if (threads <= 0)
        threads = omp_get_max_threads();
    #pragma omp parallel num_threads(threads)
    {
        orc_point* copy = NULL;
        if (faithful_copy)
            copy = (orc_point*)malloc(sizeof(orc_point) * (num_points ? num_points : 1));
        #pragma omp for schedule(dynamic)
        for (long i = 0; i < (long)num_particles; ++i) {
            float L[12];
            orc_pose_matrix(particles[i].first, tf, trig_mode, L);
            const orc_point* pts = points;
            if (faithful_copy) {
                memcpy(copy, points, sizeof(orc_point) * num_points);
                pts = copy;
            }
            orc_eval_pose(m, L, pts, num_points, prm, inv_mr, mr2, &w32[i], &w64[i]);
        }
        free(copy);
    }
}

ORC_EXPORT void orc_evaluate(const orc_coef* c, const int32_t* grid_occ, const float* data,
                             const orc_particle* particles, size_t num_particles,
                             const orc_point* points, size_t num_points, const float tf[16],
                             const orc_params* prm, int trig_mode, int threads, int faithful_copy,
                             float* w32, double* w64)
{
    orc_map m = {c, grid_occ, data, NULL, NULL, NULL, 0};
    orc_evaluate_m(&m, particles, num_particles, points, num_points, tf, prm, trig_mode, threads, faithful_copy,
                   w32, w64);
}

/* The same on a map of the 64-bit form: int64 block table, int16 TSDF
 * millimetres with truncation t read through lut (likelihood of mm at
 * mm + t - 1, 2t - 1 entries). */
ORC_EXPORT void orc_evaluate64(const orc_coef* c, const int64_t* grid_occ64, const int16_t* tsdf_mm,
                               const float* lut, int truncation_mm, const orc_particle* particles,
                               size_t num_particles, const orc_point* points, size_t num_points, const float tf[16],
                               const orc_params* prm, int trig_mode, int threads, float* w32, double* w64)
{
    orc_map m = {c, NULL, NULL, grid_occ64, tsdf_mm, lut, truncation_mm};
    orc_evaluate_m(&m, particles, num_particles, points, num_points, tf, prm, trig_mode, threads, 0, w32, w64);
}

/* Two trig bindings compared on the same inputs: the number of (particle,
 * point) pairs whose voxel index differs between trig modes a and b, and the
 * largest relative difference of a particle's float64 weight sum.  A
 * measurement of how much the choice of sin/cos matters, not a restatement. */
ORC_EXPORT void orc_trig_binding_diff(const orc_coef* c, const int32_t* grid_occ, const float* data,
                                      const orc_particle* particles, size_t num_particles,
                                      const orc_point* points, size_t num_points, const float tf[16],
                                      const orc_params* prm, int mode_a, int mode_b,
                                      uint64_t* index_mismatches, uint64_t* matrix_mismatches,
                                      double* max_rel_weight)
{
    float inv_mr, mr2;
    orc_derived(prm, &inv_mr, &mr2);
    uint64_t bad = 0, badm = 0;
    double worst = 0.0;
    #pragma omp parallel for schedule(dynamic) reduction(+:bad, badm) reduction(max:worst)
    for (long i = 0; i < (long)num_particles; ++i) {
        float La[12], Lb[12];
        orc_pose_matrix(particles[i].first, tf, mode_a, La);
        orc_pose_matrix(particles[i].first, tf, mode_b, Lb);
        if (memcmp(La, Lb, sizeof(La)) != 0) {
            ++badm;
            for (size_t j = 0; j < num_points; ++j) {
                float x = points[j].x, y = points[j].y, z = points[j].z;
                uint64_t ia = orc_get_index(c, grid_occ, La[0] * x + La[1] * y + La[2] * z + La[3],
                                            La[4] * x + La[5] * y + La[6] * z + La[7],
                                            La[8] * x + La[9] * y + La[10] * z + La[11]);
                uint64_t ib = orc_get_index(c, grid_occ, Lb[0] * x + Lb[1] * y + Lb[2] * z + Lb[3],
                                            Lb[4] * x + Lb[5] * y + Lb[6] * z + Lb[7],
                                            Lb[8] * x + Lb[9] * y + Lb[10] * z + Lb[11]);
                bad += ia != ib;
            }
            float wa32, wb32;
            double wa, wb;
            orc_map m = {c, grid_occ, data, NULL, NULL, NULL, 0};
            orc_eval_pose(&m, La, points, num_points, prm, inv_mr, mr2, &wa32, &wa);
            orc_eval_pose(&m, Lb, points, num_points, prm, inv_mr, mr2, &wb32, &wb);
            double rel = wa != 0.0 ? fabs(wa - wb) / fabs(wa) : (wb != 0.0 ? 1.0 : 0.0);
            if (rel > worst) worst = rel;
        }
    }
    *index_mismatches = bad;
    *matrix_mismatches = badm;
    *max_rel_weight = worst;
}

/* Voxel index of every (particle, point) pair, row-major [particle][point]:
 * what getIndex returns for the transformed point (data_size = rejected). */
static void orc_trace_m(const orc_map* m, const orc_particle* particles, size_t num_particles,
                        const orc_point* points, size_t num_points, const float tf[16], int trig_mode, uint64_t* out)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)num_particles; ++i) {
        float L[12];
        orc_pose_matrix(particles[i].first, tf, trig_mode, L);
        for (size_t j = 0; j < num_points; ++j) {
            float x = points[j].x, y = points[j].y, z = points[j].z;
            float tx = L[0] * x + L[1] * y + L[2] * z + L[3];
            float ty = L[4] * x + L[5] * y + L[6] * z + L[7];
            float tz = L[8] * x + L[9] * y + L[10] * z + L[11];
            out[(size_t)i * num_points + j] = orc_index_m(m, tx, ty, tz);
        }
    }
}

ORC_EXPORT void orc_trace_indices(const orc_coef* c, const int32_t* grid_occ,
                                  const orc_particle* particles, size_t num_particles,
                                  const orc_point* points, size_t num_points, const float tf[16],
                                  int trig_mode, uint64_t* out)
{
    orc_map m = {c, grid_occ, NULL, NULL, NULL, NULL, 0};
    orc_trace_m(&m, particles, num_particles, points, num_points, tf, trig_mode, out);
}

ORC_EXPORT void orc_trace_indices64(const orc_coef* c, const int64_t* grid_occ64,
                                    const orc_particle* particles, size_t num_particles,
                                    const orc_point* points, size_t num_points, const float tf[16],
                                    int trig_mode, uint64_t* out)
{
    orc_map m = {c, NULL, NULL, grid_occ64, NULL, NULL, 0};
    orc_trace_m(&m, particles, num_particles, points, num_points, tf, trig_mode, out);
}

/* Normalisation and pose estimate (:159-221).  The reference reduces with an
 * OpenMP reduction in an unspecified order; the oracle sums in index order.
 *   weights: in raw weights, out normalised weights (w / sum, float division
 *            by the float sum), like particles[i].second.
 *   sum_mode 0: float sequential sum (reference type); 1: float64 sum rounded
 *            to float.
 *   pose_out[6]: x, y, z, roll, pitch, yaw (weighted mean; circular mean of
 *            the angles via atan2 of the weighted sin/cos sums, in double).
 * Returns -1 when the sum is zero ("No particle is valid!"), else 0. */
ORC_EXPORT int orc_normalize(const orc_particle* particles, size_t n, float* weights,
                             int sum_mode, int trig_mode, double pose_out[6])
{
    float sum32 = 0.0f;
    double sum64 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        sum32 += weights[i];
        sum64 += (double)weights[i];
    }
    float weight_sum = sum_mode ? (float)sum64 : sum32;
    if (weight_sum == 0.0f)
        return -1;
    double ax = 0, ay = 0, az = 0, sr = 0, cr = 0, sp = 0, cp = 0, sy = 0, cy = 0;
    for (size_t i = 0; i < n; ++i) {
        weights[i] /= weight_sum;
        double w = (double)weights[i];
        ax += (double)particles[i].first[0] * w;
        ay += (double)particles[i].first[1] * w;
        az += (double)particles[i].first[2] * w;
        float s, c;
        orc_trig(particles[i].first[3], trig_mode, &s, &c);
        sr += (double)s * w; cr += (double)c * w;
        orc_trig(particles[i].first[4], trig_mode, &s, &c);
        sp += (double)s * w; cp += (double)c * w;
        orc_trig(particles[i].first[5], trig_mode, &s, &c);
        sy += (double)s * w; cy += (double)c * w;
    }
    pose_out[0] = ax;
    pose_out[1] = ay;
    pose_out[2] = az;
    pose_out[3] = atan2(sr, cr);
    pose_out[4] = atan2(sp, cp);
    pose_out[5] = atan2(sy, cy);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SystematicResampler::resample (novel_resampling.h:41-72)                 */
/* ------------------------------------------------------------------------ */
/* inverse_M = 1.0/M in double; U is a float drawn from [0, inverse_M) (here
 * supplied by the caller); s is the double running sum of the float weights;
 * U += inverse_M is evaluated in double and stored back into the float U.
 * Writes the source index of each emitted particle into out_idx (at most
 * capacity entries) and returns the number emitted, which can differ from M;
 * returns -1 if capacity is exceeded. */
ORC_EXPORT long orc_resample_systematic(const float* weights, size_t M, float U0,
                                        int32_t* out_idx, size_t capacity)
{
    double inverse_M = 1.0 / (double)M;
    float U = U0;
    double s = 0.0;
    size_t count = 0;
    for (size_t m = 0; m < M; ++m) {
        s += (double)weights[m];
        while (s > (double)U) {
            if (count >= capacity)
                return -1;
            out_idx[count++] = (int32_t)m;
            U = (float)((double)U + inverse_M);
        }
    }
    return (long)count;
}

/* The U recurrence alone: U_0 = U0, U_{j+1} = (float)((double)U_j + 1.0/M). */
ORC_EXPORT void orc_systematic_u(size_t M, float U0, float* out, size_t n)
{
    double inverse_M = 1.0 / (double)M;
    float U = U0;
    for (size_t j = 0; j < n; ++j) {
        out[j] = U;
        U = (float)((double)U + inverse_M);
    }
}

/* The running sum alone: s_m = s_{m-1} + (double)w_m, s_{-1} = 0. */
ORC_EXPORT void orc_cumsum(const float* weights, size_t M, double* out)
{
    double s = 0.0;
    for (size_t m = 0; m < M; ++m) {
        s += (double)weights[m];
        out[m] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Motion update: ParticleCloud::apply_model (src/particle_cloud.cpp:499-617)
 * with getAngleFromMat (src/util/util.cpp:86-111).  The six draws per particle
 * come from the library's counter-based stream (not from the reference's
 * std::mt19937, which no parallel implementation reproduces): SplitMix64
 * (Steele, Lea, Flood, OOPSLA 2014) mixes (seed, particle, draw, half), 53 bits
 * give a uniform in (0, 1), Box-Muller's cosine branch a standard normal, and
 * draw = z * stddev + mean as libstdc++'s normal_distribution forms it.
 * Parity: every float function (sinf/cosf of the particle angles, and
 * asinf/atan2f/cosf in getAngleFromMat) is the C library's here and its
 * restatement on the device, bit for bit; the double sin/cos of the draws and
 * the log/sqrt/cos of Box-Muller are each side's math library (rounded to
 * float before they meet the pose, which hides almost every last-bit
 * difference).  Pinned by closed forms (pure translation, pure yaw, rotation
 * composition) in the tests. */
/* ------------------------------------------------------------------------ */
static uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

ORC_EXPORT double orc_motion_normal(uint64_t seed, uint64_t particle, uint32_t draw)
{
    uint64_t base = orc_splitmix64(seed ^ orc_splitmix64(particle));
    uint64_t h0 = orc_splitmix64(base + (uint64_t)(2u * draw));
    uint64_t h1 = orc_splitmix64(base + (uint64_t)(2u * draw + 1u));
    double u = ((double)(h0 >> 11) + 0.5) * 0x1.0p-53;
    double v = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

ORC_EXPORT void orc_motion(const orc_particle* in, size_t n, const double mean[6], const double stddev[6],
                           uint64_t seed, int trig_mode, orc_particle* out)
{
    for (size_t i = 0; i < n; ++i) {
        const float* last = in[i].first;
        double d[6];
        for (int k = 0; k < 6; ++k)
            d[k] = orc_motion_normal(seed, i, (uint32_t)k) * stddev[k] + mean[k];
        /* :519-545: FLOAT_T sin_roll = sin(roll) with roll a double */
        float sin_roll = (float)sin(d[3]), cos_roll = (float)cos(d[3]);
        float sin_pitch = (float)sin(d[4]), cos_pitch = (float)cos(d[4]);
        float sin_yaw = (float)sin(d[5]), cos_yaw = (float)cos(d[5]);
        float tf_odom[16];
        tf_odom[0] = cos_pitch * cos_yaw;
        tf_odom[4] = cos_pitch * sin_yaw;
        tf_odom[8] = -sin_pitch;
        tf_odom[3] = (float)d[0];
        tf_odom[1] = sin_roll * sin_pitch * cos_yaw - cos_roll * sin_yaw;
        tf_odom[5] = sin_roll * sin_pitch * sin_yaw + cos_roll * cos_yaw;
        tf_odom[9] = sin_roll * cos_pitch;
        tf_odom[7] = (float)d[1];
        tf_odom[2] = cos_roll * sin_pitch * cos_yaw + sin_roll * sin_yaw;
        tf_odom[6] = cos_roll * sin_pitch * sin_yaw - sin_roll * cos_yaw;
        tf_odom[10] = cos_roll * cos_pitch;
        tf_odom[11] = (float)d[2];
        /* :549-581: particle rotation, no translation */
        float sin_alpha, cos_alpha, sin_beta, cos_beta, sin_gamma, cos_gamma;
        orc_trig(last[3], trig_mode, &sin_alpha, &cos_alpha);
        orc_trig(last[4], trig_mode, &sin_beta, &cos_beta);
        orc_trig(last[5], trig_mode, &sin_gamma, &cos_gamma);
        float tp[16];
        tp[0] = cos_beta * cos_gamma;
        tp[4] = cos_beta * sin_gamma;
        tp[8] = -sin_beta;
        tp[3] = 0;
        tp[1] = sin_alpha * sin_beta * cos_gamma - cos_alpha * sin_gamma;
        tp[5] = sin_alpha * sin_beta * sin_gamma + cos_alpha * cos_gamma;
        tp[9] = sin_alpha * cos_beta;
        tp[7] = 0;
        tp[2] = cos_alpha * sin_beta * cos_gamma + sin_alpha * sin_gamma;
        tp[6] = cos_alpha * sin_beta * sin_gamma - sin_alpha * cos_gamma;
        tp[10] = cos_alpha * cos_beta;
        tp[11] = 0;
        /* :585-600 */
        float tf[16];
        for (int r = 0; r < 3; ++r) {
            tf[4 * r + 0] = tp[4 * r] * tf_odom[0] + tp[4 * r + 1] * tf_odom[4] + tp[4 * r + 2] * tf_odom[8];
            tf[4 * r + 1] = tp[4 * r] * tf_odom[1] + tp[4 * r + 1] * tf_odom[5] + tp[4 * r + 2] * tf_odom[9];
            tf[4 * r + 2] = tp[4 * r] * tf_odom[2] + tp[4 * r + 1] * tf_odom[6] + tp[4 * r + 2] * tf_odom[10];
            tf[4 * r + 3] = tp[4 * r] * tf_odom[3] + tp[4 * r + 1] * tf_odom[7] + tp[4 * r + 2] * tf_odom[11] +
                            tp[4 * r + 3];
        }
        /* getAngleFromMat (util.cpp:86-111).  Its arguments are floats: in
         * the C++ reference with libstdc++, fabs, atan2, asin and cos bind to
         * the float functions (fabsf, atan2f, asinf, cosf) and the divisions
         * are float; mode 2 keeps the C binding (double functions). */
        float roll, pitch, yaw;
        if (trig_mode == 2) {
            if (fabs(tf[8]) >= 1) {
                yaw = 0;
                double delta = atan2(tf[9], tf[10]);
                if (tf[8] < 0) {
                    pitch = M_PI / 2.0;
                    roll = delta;
                } else {
                    pitch = -M_PI / 2.0;
                    roll = delta;
                }
            } else {
                pitch = -asin(tf[8]);
                roll = atan2(tf[9] / cos(pitch), tf[10] / cos(pitch));
                yaw = atan2(tf[4] / cos(pitch), tf[0] / cos(pitch));
            }
        } else {
            volatile float m8 = tf[8], m9 = tf[9], m10 = tf[10];
            if (fabsf(m8) >= 1) {
                yaw = 0;
                float delta = atan2f(m9, m10);
                if (tf[8] < 0) {
                    pitch = M_PI / 2.0;
                    roll = delta;
                } else {
                    pitch = -M_PI / 2.0;
                    roll = delta;
                }
            } else {
                pitch = -asinf(m8);
                volatile float pv = pitch;
                float cp = cosf(pv);
                roll = atan2f(tf[9] / cp, tf[10] / cp);
                yaw = atan2f(tf[4] / cp, tf[0] / cp);
            }
        }
        out[i].first[0] = last[0] + tf[3];
        out[i].first[1] = last[1] + tf[7];
        out[i].first[2] = last[2] + tf[11];
        out[i].first[3] = roll;
        out[i].first[4] = pitch;
        out[i].first[5] = yaw;
        out[i].second = in[i].second;
    }
}

/* ------------------------------------------------------------------------ */
/* Scan reduction of TSDFEvaluator::evaluateParticles
 * (src/evaluation/tsdf_evaluator.cpp:304-376):
 *   - points with sqrtf(x^2 + y^2 + z^2) < 1 are dropped; the ring iterator
 *     and the running index advance only for kept points (:319-331), so the
 *     k-th kept point reads ring[k] and gets index k;
 *   - centre = floor(p / cell) * cell + cell / 2 per axis (float);
 *   - an unordered_set keyed by (ring, centre) keeps the first point
 *     inserted per key (float equality; NaN never equal);
 *   - output grouped by ring 0..63, each ring in index order, the original
 *     point of each kept entry.
 * Returns the number of points written, or -1 if a kept point's ring is
 * outside [0, 64) (the reference indexes a 64-entry vector with it). */
/* ------------------------------------------------------------------------ */
typedef struct {
    int ring;
    uint32_t index;
    float c[3];
    float orig[3];
} orc_red;

static int orc_red_cmp_key(const void* pa, const void* pb)
{
    const orc_red* a = (const orc_red*)pa;
    const orc_red* b = (const orc_red*)pb;
    if (a->ring != b->ring) return a->ring < b->ring ? -1 : 1;
    for (int k = 0; k < 3; ++k) {
        uint32_t ua, ub;
        memcpy(&ua, &a->c[k], 4);
        memcpy(&ub, &b->c[k], 4);
        if (ua != ub) return ua < ub ? -1 : 1;
    }
    return a->index < b->index ? -1 : (a->index > b->index ? 1 : 0);
}

static int orc_red_cmp_out(const void* pa, const void* pb)
{
    const orc_red* a = (const orc_red*)pa;
    const orc_red* b = (const orc_red*)pb;
    if (a->ring != b->ring) return a->ring < b->ring ? -1 : 1;
    return a->index < b->index ? -1 : (a->index > b->index ? 1 : 0);
}

ORC_EXPORT long orc_reduce_scan(const float* xyz, const int16_t* ring, size_t n, float cell, float* out_xyz)
{
    const float half = cell / 2;
    orc_red* v = (orc_red*)malloc(sizeof(orc_red) * (n ? n : 1));
    size_t kept = 0;
    for (size_t j = 0; j < n; ++j) {
        float x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
        float dist = sqrtf(x * x + y * y + z * z);
        if (dist < 1.0)
            continue;
        orc_red r;
        r.ring = ring[kept]; /* the ring iterator lags behind skipped points */
        r.index = (uint32_t)kept;
        r.c[0] = floorf(x / cell) * cell + half;
        r.c[1] = floorf(y / cell) * cell + half;
        r.c[2] = floorf(z / cell) * cell + half;
        r.orig[0] = x;
        r.orig[1] = y;
        r.orig[2] = z;
        if (r.ring < 0 || r.ring >= 64) {
            free(v);
            return -1;
        }
        v[kept++] = r;
    }
    qsort(v, kept, sizeof(orc_red), orc_red_cmp_key);
    size_t w = 0;
    for (size_t i = 0; i < kept; ++i) {
        int nan = v[i].c[0] != v[i].c[0] || v[i].c[1] != v[i].c[1] || v[i].c[2] != v[i].c[2];
        int first = i == 0 || nan || v[i].ring != v[i - 1].ring || v[i].c[0] != v[i - 1].c[0] ||
                    v[i].c[1] != v[i - 1].c[1] || v[i].c[2] != v[i - 1].c[2];
        if (first)
            v[w++] = v[i];
    }
    qsort(v, w, sizeof(orc_red), orc_red_cmp_out);
    for (size_t i = 0; i < w; ++i) {
        out_xyz[3 * i] = v[i].orig[0];
        out_xyz[3 * i + 1] = v[i].orig[1];
        out_xyz[3 * i + 2] = v[i].orig[2];
    }
    free(v);
    return (long)w;
}

/* ------------------------------------------------------------------------ */
/* CudaEvaluator::evaluate(particles, PointCloud2, tf) point reduction      */
/* (src/cuda/cuda_evaluator.cu:78-116):                                      */
/*   - every point is kept (no range test), read with its ring as int;     */
/*   - centre = floor(v / 0.064) * 0.064 + 0.032 per axis, in double (the  */
/*     literals are double) and stored as float (:95-97);                  */
/*   - an unordered_set<CudaPoint> keeps each distinct centre once (float  */
/*     equality; NaN never equal), and the reduced scan is the centres.    */
/* The set's iteration order is unspecified; the order written here is    */
/* the library's documented one: by the ring, then the index, of each      */
/* centre's lowest-index point.                                            */
/* Returns the number of centres written.                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t ring;
    uint32_t index;
    float c[3];
} orc_cloud;

static int orc_cloud_cmp_key(const void* pa, const void* pb)
{
    const orc_cloud* a = (const orc_cloud*)pa;
    const orc_cloud* b = (const orc_cloud*)pb;
    for (int k = 0; k < 3; ++k) {
        uint32_t ua, ub;
        memcpy(&ua, &a->c[k], 4);
        memcpy(&ub, &b->c[k], 4);
        if (ua != ub) return ua < ub ? -1 : 1;
    }
    return a->index < b->index ? -1 : (a->index > b->index ? 1 : 0);
}

static int orc_cloud_cmp_out(const void* pa, const void* pb)
{
    const orc_cloud* a = (const orc_cloud*)pa;
    const orc_cloud* b = (const orc_cloud*)pb;
    if (a->ring != b->ring) return a->ring < b->ring ? -1 : 1;
    return a->index < b->index ? -1 : (a->index > b->index ? 1 : 0);
}

ORC_EXPORT long orc_reduce_cloud(const float* xyz, const int32_t* ring, size_t n, float* out_xyz)
{
    orc_cloud* v = (orc_cloud*)malloc(sizeof(orc_cloud) * (n ? n : 1));
    for (size_t j = 0; j < n; ++j) {
        v[j].ring = ring[j];
        v[j].index = (uint32_t)j;
        for (int k = 0; k < 3; ++k)
            v[j].c[k] = (float)(floor((double)xyz[3 * j + k] / 0.064) * 0.064 + 0.032);
    }
    qsort(v, n, sizeof(orc_cloud), orc_cloud_cmp_key);
    size_t w = 0;
    for (size_t i = 0; i < n; ++i) {
        int nan = v[i].c[0] != v[i].c[0] || v[i].c[1] != v[i].c[1] || v[i].c[2] != v[i].c[2];
        int first = i == 0 || nan || v[i].c[0] != v[i - 1].c[0] || v[i].c[1] != v[i - 1].c[1] ||
                    v[i].c[2] != v[i - 1].c[2];
        if (first)
            v[w++] = v[i];
    }
    qsort(v, w, sizeof(orc_cloud), orc_cloud_cmp_out);
    for (size_t i = 0; i < w; ++i)
        for (int k = 0; k < 3; ++k)
            out_xyz[3 * i + k] = v[i].c[k];
    free(v);
    return (long)w;
}
