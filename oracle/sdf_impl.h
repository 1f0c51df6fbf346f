/*
 * Signed-distance objects of the oracle (test infrastructure only), written
 * once for a real type R and instantiated by the includer:
 *
 *   R, SR(name), SQRT, FLOOR, FABS    the arithmetic of the instance
 *
 * trace.c instantiates it in float (the traversal oracle, bit-exact with the
 * GPU march), render_impl.h in float and double (the integrator oracles).
 *
 * Restated from core/object/sdf:
 *   SR(sdf_torus)     SdfTorus (sdf_torus.cuh:14-18)
 *   SR(sdf_volume)    SampleSDFAdaptive01 = nearest voxel when its value is
 *                     farther than voxelSizeMax_, else trilinear
 *                     (sdf_volume.cuh:75-160), queried at u + 0.5
 *   SR(sdf_unit)      the unit-space distance of the object's shape
 *   SR(sdf_march)     RaymarchLocal (sdf_base.cuh:54-116): sphere tracing in
 *                     object space from max(0, tEnter) to tExit, at most 512
 *                     steps (SDF_STEPS), step max(|d| * s, eps / 2) with s = 0.9
 *                     for volumes (SDF_VOLUME_STEP_SCALE) and 1 otherwise; a
 *                     ray that starts inside (d <= 0 at its origin) reports the
 *                     first point with 0 < d < eps, one that starts outside
 *                     the first with -eps < d < 0
 *   SR(sdf_normal)    sdfNormalUnit + UnitNormalToLocal (sdf_base.cuh:26-51):
 *                     central differences of step eps / sMin in unit space
 *
 * Arithmetic order is the GPU's (csrc/objects.hpp): no contraction, dot
 * products x + y then + z, IEEE division and square root.
 */

static R SR(sdf_fmax)(R a, R b) { return (R)fmax((double)a, (double)b); }
static R SR(sdf_fmin)(R a, R b) { return (R)fmin((double)a, (double)b); }

/* SdfTorus: q = (|u.xz| - R, u.y), |q| - r */
static R SR(sdf_torus)(const orc_object* ob, R ux, R uy, R uz)
{
    const R qx = SQRT(ux * ux + uz * uz) - (R)ob->sdf[0];
    return SQRT(qx * qx + uy * uy) - (R)ob->sdf[1];
}

static R SR(sdf_voxel)(const orc_object* ob, int x, int y, int z)
{
    return (R)ob->vol[((size_t)z * ob->dims[1] + (size_t)y) * ob->dims[0] + (size_t)x];
}

static int SR(sdf_clampi)(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

/* SampleSDFAdaptive01 at p01 = u + 0.5 (sdf_volume.cuh:81-160) */
static R SR(sdf_volume)(const orc_object* ob, R ux, R uy, R uz)
{
    const int nx = (int)ob->dims[0], ny = (int)ob->dims[1], nz = (int)ob->dims[2];
    const R gx = (ux + (R)0.5) * (R)(nx - 1), gy = (uy + (R)0.5) * (R)(ny - 1), gz = (uz + (R)0.5) * (R)(nz - 1);
    const int xi = SR(sdf_clampi)((int)FLOOR(gx + (R)0.5), nx - 1);
    const int yi = SR(sdf_clampi)((int)FLOOR(gy + (R)0.5), ny - 1);
    const int zi = SR(sdf_clampi)((int)FLOOR(gz + (R)0.5), nz - 1);
    const R dn = SR(sdf_voxel)(ob, xi, yi, zi);
    if (FABS(dn) > (R)ob->vmax) return dn;
    int x0 = (int)FLOOR(gx), y0 = (int)FLOOR(gy), z0 = (int)FLOOR(gz);
    const int x1 = SR(sdf_clampi)(x0 + 1, nx - 1), y1 = SR(sdf_clampi)(y0 + 1, ny - 1),
              z1 = SR(sdf_clampi)(z0 + 1, nz - 1);
    x0 = SR(sdf_clampi)(x0, nx - 1); y0 = SR(sdf_clampi)(y0, ny - 1); z0 = SR(sdf_clampi)(z0, nz - 1);
    const R fx = SR(sdf_fmin)(SR(sdf_fmax)(gx - (R)x0, (R)0), (R)1);
    const R fy = SR(sdf_fmin)(SR(sdf_fmax)(gy - (R)y0, (R)0), (R)1);
    const R fz = SR(sdf_fmin)(SR(sdf_fmax)(gz - (R)z0, (R)0), (R)1);
    const R c00 = SR(sdf_voxel)(ob, x0, y0, z0) * ((R)1 - fx) + SR(sdf_voxel)(ob, x1, y0, z0) * fx;
    const R c10 = SR(sdf_voxel)(ob, x0, y1, z0) * ((R)1 - fx) + SR(sdf_voxel)(ob, x1, y1, z0) * fx;
    const R c01 = SR(sdf_voxel)(ob, x0, y0, z1) * ((R)1 - fx) + SR(sdf_voxel)(ob, x1, y0, z1) * fx;
    const R c11 = SR(sdf_voxel)(ob, x0, y1, z1) * ((R)1 - fx) + SR(sdf_voxel)(ob, x1, y1, z1) * fx;
    const R c0 = c00 * ((R)1 - fy) + c10 * fy;
    const R c1 = c01 * ((R)1 - fy) + c11 * fy;
    return c0 * ((R)1 - fz) + c1 * fz;
}

static R SR(sdf_unit)(const orc_object* ob, R ux, R uy, R uz)
{
    return ob->type == ORC_OBJ_SDF_TORUS ? SR(sdf_torus)(ob, ux, uy, uz) : SR(sdf_volume)(ob, ux, uy, uz);
}

/* ToUnitSpace (sdf_base.cuh:12-21): [-extent, extent] -> [-0.5, 0.5] */
static void SR(sdf_to_unit)(const orc_object* ob, const R l[3], R u[3])
{
    for (int a = 0; a < 3; ++a) u[a] = l[a] / ((R)2 * SR(sdf_fmax)((R)ob->extent[a], (R)1e-8f));
}

/* LocalScaleMin (sdf_base.cuh:23-26) */
static R SR(sdf_smin)(const orc_object* ob)
{
    return (R)2 * SR(sdf_fmin)((R)ob->extent[0], SR(sdf_fmin)((R)ob->extent[1], (R)ob->extent[2]));
}

/* RaymarchLocal: 1 with the hit distance on a crossing inside [tEnter, tExit] */
static int SR(sdf_march)(const orc_object* ob, const R lo[3], const R ld[3], R t_enter, R t_exit, R eps, R* t_hit,
                         int* inside)
{
    const R min_step = eps * (R)0.5;
    const R step_scale = ob->type == ORC_OBJ_SDF_VOLUME ? (R)0.9f : (R)1;
    const R smin = SR(sdf_smin)(ob);
    R u[3];
    SR(sdf_to_unit)(ob, lo, u);
    const int start_in = SR(sdf_unit)(ob, u[0], u[1], u[2]) <= (R)0;
    *inside = start_in;
    R t = SR(sdf_fmax)((R)0, t_enter);
    for (int i = 0; i < 512; ++i) {
        const R tl = t;
        const R p[3] = { lo[0] + ld[0] * t, lo[1] + ld[1] * t, lo[2] + ld[2] * t };
        SR(sdf_to_unit)(ob, p, u);
        const R dl = SR(sdf_unit)(ob, u[0], u[1], u[2]) * smin;
        const R ad = FABS(dl);
        if (start_in ? (dl > (R)0 && dl < eps) : (dl < (R)0 && -dl < eps)) {
            *t_hit = tl;
            return 1;
        }
        t = t + SR(sdf_fmax)(ad * step_scale, min_step);
        if (t > t_exit) break;
    }
    return 0;
}

/* sdfNormalUnit + UnitNormalToLocal: the object-space normal at local point lp */
static void SR(sdf_normal)(const orc_object* ob, const R lp[3], R eps, R n[3])
{
    R u[3];
    SR(sdf_to_unit)(ob, lp, u);
    const R h = eps / SR(sdf_fmax)(SR(sdf_smin)(ob), (R)1e-8f);
    const R dx = SR(sdf_unit)(ob, u[0] + h, u[1], u[2]) - SR(sdf_unit)(ob, u[0] - h, u[1], u[2]);
    const R dy = SR(sdf_unit)(ob, u[0], u[1] + h, u[2]) - SR(sdf_unit)(ob, u[0], u[1] - h, u[2]);
    const R dz = SR(sdf_unit)(ob, u[0], u[1], u[2] + h) - SR(sdf_unit)(ob, u[0], u[1], u[2] - h);
    R g[3] = { dx, dy, dz };
    for (int pass = 0; pass < 2; ++pass) {
        const R l2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        if (l2 <= (R)1e-20f) {
            g[0] = g[1] = g[2] = 0;
        } else {
            const R inv = (R)1 / SQRT(l2);
            g[0] = g[0] * inv; g[1] = g[1] * inv; g[2] = g[2] * inv;
        }
        if (pass == 0)
            for (int a = 0; a < 3; ++a) g[a] = g[a] / ((R)2 * SR(sdf_fmax)((R)ob->extent[a], (R)1e-8f));
    }
    n[0] = g[0]; n[1] = g[1]; n[2] = g[2];
}
