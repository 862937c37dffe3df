/*
 * jetshow.cpp -- the reference's class API over the C ABI.  Every compute
 * call goes to the GPU through include/jetshow_amd.h; nothing here computes
 * radiative transfer on the CPU.
 */
#include "jetshow.h"

#include <cmath>
#include <cstring>
#include <mutex>

namespace jetshow {

static void set3(double *dst, const Vector3d &v) { dst[0] = v[0]; dst[1] = v[1]; dst[2] = v[2]; }

/* ---- Ray (src/Ray.cpp:9-19) ---- */
Ray::Ray(const Vector3d &origin, const Vector3d &direction) : origin_(origin), direction_(direction) {
    double z = (direction_[0] * direction_[0] + direction_[1] * direction_[1]) + direction_[2] * direction_[2];
    if (z > 0.0) {
        double r = std::sqrt(z);
        for (int i = 0; i < 3; ++i) direction_[i] /= r;
    }
}

Vector3d Ray::point(double t) const {
    return Vector3d(origin_[0] + direction_[0] * t, origin_[1] + direction_[1] * t, origin_[2] + direction_[2] * t);
}

/* ---- Context ---- */
Context::Context(int device, size_t hbm_budget) {
    int rc = jsx_ctx_create(device, hbm_budget, &ctx_);
    if (rc) throw Error(rc, "jsx_ctx_create failed: no usable GPU");
}
Context::~Context() {
    if (ctx_) jsx_ctx_destroy(ctx_);
}
void Context::check(int rc) const {
    if (rc) throw Error(rc, jsx_last_error(ctx_));
}
Context &Context::instance() {
    static Context ctx(0, 0);
    return ctx;
}
int32_t Context::upload_hull(const std::vector<double> &planes) {
    if (planes.size() % 4) throw Error(JSX_EINVAL, "hull planes come in fours");
    int32_t h = 0;
    check(jsx_hull_upload(ctx_, planes.data(), (int32_t)(planes.size() / 4), &h));
    return h;
}

int32_t Context::upload_points(const std::vector<double> &zrv, double nan_value) {
    if (zrv.size() % 3) throw Error(JSX_EINVAL, "nodes come in threes (z, r, value)");
    int32_t h = 0;
    const int rc = jsx_grid_upload_points(ctx_, zrv.data(), (int64_t)(zrv.size() / 3), nan_value, &h);
    if (rc == JSX_EUNSUPPORTED) throw Unsupported("the nodes span no triangle (fewer than three, or all on one line)");
    check(rc);
    return h;
}

int32_t Context::upload_grid(const std::vector<double> &values, int nz, int nr, double z0, double dz, double r0,
                             double dr, double nan_value) {
    if ((size_t)nz * (size_t)nr != values.size()) throw Error(JSX_EINVAL, "grid size mismatch");
    jsx_grid_desc g;
    g.nz = nz; g.nr = nr; g.z0 = z0; g.dz = dz; g.r0 = r0; g.dr = dr; g.nan_value = nan_value;
    int32_t h = 0;
    check(jsx_grid_upload(ctx_, &g, values.data(), &h));
    return h;
}

std::vector<double> Context::fields(const jsx_jet_desc &d, const std::vector<double> &points,
                                    const std::vector<double> &n_los, double nu, uint32_t flags) {
    if (points.size() % 3 || points.size() != n_los.size()) throw Error(JSX_EINVAL, "points and n_los: n x 3 each");
    const int64_t n = (int64_t)(points.size() / 3);
    std::vector<double> out((size_t)n * JSX_FIELDS_STRIDE);
    if (n == 0) return out;
    jsx_solver_params p;
    std::memset(&p, 0, sizeof p);
    p.flags = flags;
    check(jsx_fields_host(ctx_, &d, &p, points.data(), n_los.data(), n, nu, out.data()));
    return out;
}

/* a descriptor with every kind set to a plain default; a class lowers itself
 * over it to be evaluated on its own */
static jsx_jet_desc default_desc() {
    jsx_jet_desc d;
    std::memset(&d, 0, sizeof d);
    d.geometry = JSX_GEOM_CONE; d.geo_axis[2] = 1.0; d.geo_param = 0.1; d.geo_big_scale = 100.0;
    d.bfield = JSX_B_RANDOM_SCALAR; d.b0 = 1.0;
    d.nfield = JSX_N_CONST; d.n0 = 1.0;
    d.vfield = JSX_V_CONST_FLAT; d.gamma0 = 2.0;
    d.s = 2.5; d.gamma_min = 100.0;
    return d;
}

/* the JSX_FIELDS_STRIDE values at one point (unit LOS along +z, 1 GHz when
 * only a field's own value is wanted) */
static std::vector<double> one_point(const jsx_jet_desc &d, const Vector3d &p, const Vector3d &n_los, double nu,
                                     uint32_t flags = 0) {
    return Context::instance().fields(d, {p[0], p[1], p[2]}, {n_los[0], n_los[1], n_los[2]}, nu, flags);
}
static std::vector<double> one_point(const jsx_jet_desc &d, const Vector3d &p) {
    return one_point(d, p, Vector3d(0, 0, 1), 1e9);
}

/* ---- frame transforms (src/utils.cpp:147-168) ---- */
static double dot3(const Vector3d &a, const Vector3d &b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
double getG(const Vector3d &v) {
    Vector3d beta(v[0] / c, v[1] / c, v[2] / c);
    return std::sqrt(1. / (1. - dot3(beta, beta)));
}
double getD(const Vector3d &n_los, const Vector3d &v) {
    Vector3d beta(v[0] / c, v[1] / c, v[2] / c);
    return 1. / (getG(v) * (1. - dot3(beta, n_los)));
}
Vector3d get_n_los_prime(const Vector3d &n_los, const Vector3d &v) {
    double df = getD(n_los, v), gamma = getG(v);
    Vector3d beta(v[0] / c, v[1] / c, v[2] / c);
    double f = (df + 1.) * (gamma / (gamma + 1.));
    return Vector3d(df * n_los[0] - f * beta[0], df * n_los[1] - f * beta[1], df * n_los[2] - f * beta[2]);
}
Vector3d getBprime(const Vector3d &b, const Vector3d &v) {
    double gamma = getG(v);
    double s1 = gamma / ((1. + gamma) * (c * c));
    double vb = dot3(v, b);
    return Vector3d(b[0] / gamma + (s1 * v[0]) * vb, b[1] / gamma + (s1 * v[1]) * vb, b[2] / gamma + (s1 * v[2]) * vb);
}

/* ---- field values (the class's bf / nf / vf on the GPU) ---- */
double ScalarBField::bf(const Vector3d &point) const {
    jsx_jet_desc d = default_desc();
    lower(d);
    return one_point(d, point)[4];
}
Vector3d VectorBField::bf(const Vector3d &point) const {
    jsx_jet_desc d = default_desc();
    lower(d);
    auto r = one_point(d, point);
    return Vector3d(r[4], r[5], r[6]);
}
Vector3d VectorBField::bf_plasma_frame(const Vector3d &point, Vector3d &v) const {
    Vector3d b = bf(point);
    return in_plasma_frame_ ? b : getBprime(b, v);
}
double NField::nf(const Vector3d &point) const {
    jsx_jet_desc d = default_desc();
    lower(d);
    return one_point(d, point)[11];
}
Vector3d VField::vf(const Vector3d &point) const {
    jsx_jet_desc d = default_desc();
    lower(d);
    auto r = one_point(d, point);
    return Vector3d(r[0], r[1], r[2]);
}
bool Geometry::is_within(Vector3d &point) const {
    jsx_jet_desc d = default_desc();
    lower(d);
    return one_point(d, point)[29] != 0.0;
}

/* ---- geometry ---- */
std::list<Intersection> BaseGeometry::hit(Ray &ray) const {
    jsx_jet_desc d = default_desc();
    lower(d);
    double r[6] = {ray.origin()[0], ray.origin()[1], ray.origin()[2],
                   ray.direction()[0], ray.direction()[1], ray.direction()[2]};
    int32_t ns = 0;
    double seg[12];
    Context &C = Context::instance();
    C.check(jsx_hit_host(C.get(), &d, nullptr, r, 1, &ns, seg));
    std::list<Intersection> out;
    for (int s = 0; s < ns; ++s)
        out.emplace_back(ray.direction(), Vector3d(seg[6 * s], seg[6 * s + 1], seg[6 * s + 2]),
                         Vector3d(seg[6 * s + 3], seg[6 * s + 4], seg[6 * s + 5]));
    return out;
}

Cone::Cone(const Vector3d &origin, const Vector3d &direction, const double &angle, const double &scale)
    : origin_(origin), direction_(direction), angle_(angle), big_scale_(scale) {}
void Cone::lower(jsx_jet_desc &d) const {
    d.geometry = JSX_GEOM_CONE;
    set3(d.geo_origin, origin_); set3(d.geo_axis, direction_);
    d.geo_param = angle_; d.geo_big_scale = big_scale_;
    d.z_clip_min = zmin_; d.z_clip_max = zmax_;
}
Parabaloid::Parabaloid(const Vector3d &origin, const Vector3d &direction, const double &r0, const double &big_scale)
    : origin_(origin), direction_(direction), r0_(r0), big_scale_(big_scale) {}
void Parabaloid::lower(jsx_jet_desc &d) const {
    d.geometry = JSX_GEOM_PARABOLOID;
    set3(d.geo_origin, origin_); set3(d.geo_axis, direction_);
    d.geo_param = r0_; d.geo_big_scale = big_scale_;
}
Cylinder::Cylinder(const Vector3d &origin, const Vector3d &direction, const double &r)
    : origin_(origin), direction_(direction), r_(r) {}
void Cylinder::lower(jsx_jet_desc &d) const {
    d.geometry = JSX_GEOM_CYLINDER;
    set3(d.geo_origin, origin_); set3(d.geo_axis, direction_);
    d.geo_param = r_; d.geo_big_scale = 100. * r_;
    d.z_clip_min = zmin_; d.z_clip_max = zmax_;
}

ParabaloidCone::ParabaloidCone(const Vector3d &origin, const Vector3d &direction, const double &r0, const double &z0,
                               const double &scale)
    : origin_(origin), direction_(direction), r0_(r0), z0_(z0), big_scale_(scale) {}
void ParabaloidCone::lower(jsx_jet_desc &d) const {
    d.geometry = JSX_GEOM_PARABOLOID_CONE;
    set3(d.geo_origin, origin_); set3(d.geo_axis, direction_);
    d.geo_param = r0_; d.geo_z0 = z0_; d.geo_big_scale = big_scale_;
}

void SimulationGeometry::lower(jsx_jet_desc &d) const {
    d.geometry = JSX_GEOM_CONVEX_HULL;
    d.geo_table = h_;
}

/* ---- fields ---- */
void RandomScalarBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_RANDOM_SCALAR; d.b0 = b_0_; d.b_exp = m_b_; d.b_in_plasma_frame = 1; }
void RandomScalarBFieldZ::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_RANDOM_SCALAR_Z; d.b0 = b_0_; d.b_exp = m_b_; d.b_in_plasma_frame = 1; }
void ToroidalBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_TOROIDAL; d.b0 = b_0_; d.b_exp = n_b_; d.b_in_plasma_frame = in_plasma_frame_; }
void ConstCylinderBFieldZ::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_CONST_CYLINDER_Z; d.b0 = b_0_; d.b_exp = n_b_; d.b_in_plasma_frame = in_plasma_frame_; }
void ToroidalRBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_TOROIDAL_R; d.b0 = b_0_; d.b_exp = n_b_; d.b_in_plasma_frame = in_plasma_frame_; }
void HelicalRBField::lower(jsx_jet_desc &d) const {
    d.bfield = JSX_B_HELICAL_R; d.b0 = b_0_; d.b_exp = n_b_; d.b_pitch = pitch_; d.b_in_plasma_frame = in_plasma_frame_;
}
void CompositeRandomScalarBFieldZ::lower(jsx_jet_desc &d) const {
    d.bfield = JSX_B_COMPOSITE_RANDOM_SCALAR_Z; d.b0 = b_0_; d.b_exp = m_in_; d.b_exp2 = m_out_; d.b_z0 = z0_;
    d.b_in_plasma_frame = 1;
}
void ConstCylinderBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_CONST_CYLINDER; d.b0 = b_0_; d.b_exp = n_b_; d.b_in_plasma_frame = in_plasma_frame_; }
void RadialConicalBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_RADIAL_CONICAL; d.b0 = b_0_; d.b_exp = n_b_; d.b_in_plasma_frame = in_plasma_frame_; }
void HelicalCylinderBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_HELICAL_CYLINDER; d.b0 = b_0_; d.b_pitch = pitch_; d.b_in_plasma_frame = in_plasma_frame_; }
void SpiralConicalBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_SPIRAL_CONICAL; d.b0 = b_0_; d.b_pitch = pitch_; d.b_in_plasma_frame = in_plasma_frame_; }
void ForceFreeCylindricalBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_FORCE_FREE_CYLINDRICAL; d.b0 = b_0_; d.b_mu = mu_; d.b_in_plasma_frame = in_plasma_frame_; }
void SimulationBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_GRID_VECTOR; d.grid_b = gp_; d.grid_b2 = gf_; d.b_in_plasma_frame = in_plasma_frame_; }
void SimulationScalarBField::lower(jsx_jet_desc &d) const { d.bfield = JSX_B_GRID_SCALAR; d.grid_b = g_; d.b_in_plasma_frame = 1; }
void ConstNField::lower(jsx_jet_desc &d) const { d.nfield = JSX_N_CONST; d.n0 = n_; d.n_in_plasma_frame = in_plasma_frame_; }
void BKNField::lower(jsx_jet_desc &d) const { d.nfield = JSX_N_BK; d.n0 = n_0_; d.n_exp = n_n_; d.n_in_plasma_frame = in_plasma_frame_; }
void BKNFieldZ::lower(jsx_jet_desc &d) const { d.nfield = JSX_N_BK_Z; d.n0 = n_0_; d.n_exp = n_n_; d.n_in_plasma_frame = in_plasma_frame_; }
void CompositeBKNField::lower(jsx_jet_desc &d) const {
    d.nfield = JSX_N_COMPOSITE_BK; d.n0 = n_0_; d.n_exp = n_in_; d.n_exp2 = n_out_; d.n_z0 = z0_;
    d.n_in_plasma_frame = in_plasma_frame_;
}
void SimulationNField::lower(jsx_jet_desc &d) const { d.nfield = JSX_N_GRID; d.grid_n = g_; d.n_in_plasma_frame = in_plasma_frame_; }
void ConstFlatVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_CONST_FLAT; d.gamma0 = gamma_; }
void ConstCentralVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_CONST_CENTRAL; d.gamma0 = gamma_; set3(d.v_origin, origin_); }
void LinearGammaZVField::lower(jsx_jet_desc &d) const {
    d.vfield = JSX_V_LINEAR_GAMMA_Z; d.gamma0 = g0_; d.gamma1 = g1_; d.gamma_z0 = z0_; d.gamma_z1 = z1_;
}
void SimulationVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_GRID; d.grid_v = g_; }
void ShearedFlatVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_SHEARED_FLAT; d.gamma0 = ga_; d.gamma1 = gb_; d.v_r = r_; }
void SheathFlatVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_SHEATH_FLAT; d.gamma0 = ga_; d.gamma1 = gb_; d.v_r = r_; }
void ShearedCentralVField::lower(jsx_jet_desc &d) const {
    d.vfield = JSX_V_SHEARED_CENTRAL; d.gamma0 = ga_; d.gamma1 = gb_; d.v_theta = theta_; set3(d.v_origin, origin_);
}
void SheathCentralVField::lower(jsx_jet_desc &d) const {
    d.vfield = JSX_V_SHEATH_CENTRAL; d.gamma0 = ga_; d.gamma1 = gb_; d.v_theta = theta_; set3(d.v_origin, origin_);
}
void ConstParabolicVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_CONST_PARABOLIC; d.gamma0 = g_; d.v_Rz0 = Rz0_; }
void AccParabolicVField::lower(jsx_jet_desc &d) const { d.vfield = JSX_V_ACC_PARABOLIC; d.gamma0 = g_; d.v_R0 = R0_; d.v_Rz0 = Rz0_; }
void ShearedAccParabolicVField::lower(jsx_jet_desc &d) const {
    d.vfield = JSX_V_SHEARED_ACC_PARABOLIC; d.gamma0 = ga_; d.gamma1 = gb_; d.v_R0 = R0_; d.v_Rz0 = Rz0_;
}
void AccParabolicConstConeVField::lower(jsx_jet_desc &d) const {
    d.vfield = JSX_V_ACC_PARABOLIC_CONST_CONE; d.gamma0 = g_; d.v_Rz0 = Rz0_; d.v_z0 = z0_;
}
void ShearedAccParabolicConstConeVField::lower(jsx_jet_desc &d) const {
    d.vfield = JSX_V_SHEARED_ACC_PARABOLIC_CONST_CONE; d.gamma0 = ga_; d.gamma1 = gb_; d.v_Rz0 = Rz0_; d.v_z0 = z0_;
}

/* ---- Jet ---- */
Jet::Jet(BaseGeometry *geo, VField *vfield, ScalarBField *bField, NField *nField)
    : geometry_(geo), vfield_(vfield), sbfield_(bField), nfield_(nField) {}
Jet::Jet(BaseGeometry *geo, VField *vfield, VectorBField *bField, NField *nField)
    : geometry_(geo), vfield_(vfield), vbfield_(bField), nfield_(nField) {}

std::list<Intersection> Jet::hit(Ray &ray) { return geometry_->hit(ray); }

/* Jet::getKI, getEtaI, getV and the restored getters: the lab-frame values
 * the kernels compute at the point (layout of jsx_fields_host) */
static double jet_value(const Jet &j, Vector3d &point, Vector3d &n_los, double nu, int k) {
    return one_point(j.descriptor(), point, n_los, nu, j.flags)[k];
}
double Jet::getKI(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 18); }
double Jet::getKQ(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 19); }
double Jet::getKU(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 20); }
double Jet::getKV(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 21); }
double Jet::getKF(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 22); }
double Jet::getKC(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 23); }
double Jet::gethQ(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 24); }
double Jet::getEtaI(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 25); }
double Jet::getEtaQ(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 26); }
double Jet::getEtaU(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 27); }
double Jet::getEtaV(Vector3d &p, Vector3d &n, double nu) { return jet_value(*this, p, n, nu, 28); }
const Vector3d Jet::getV(const Vector3d &point) {
    auto r = one_point(descriptor(), point);
    return Vector3d(r[0], r[1], r[2]);
}

jsx_jet_desc Jet::descriptor() const {
    jsx_jet_desc d;
    std::memset(&d, 0, sizeof d);
    geometry_->lower(d);
    if (sbfield_) sbfield_->lower(d);
    if (vbfield_) vbfield_->lower(d);
    nfield_->lower(d);
    vfield_->lower(d);
    d.s = s_;
    d.gamma_min = gamma_min_;
    d.thermal_fraction = thermal_;
    return d;
}

/* ---- ImagePlane ---- */
ImagePlane::ImagePlane(std::pair<unsigned long int, unsigned long int> image_size_, double pixel_size_start,
                       double pixel_size_stop, double los_angle)
    : image_size(image_size_) {
    desc_.n_across = (int64_t)image_size_.first;
    desc_.n_along = (int64_t)image_size_.second;
    desc_.lg_pixel_size_start_cm = pixel_size_start;
    desc_.lg_pixel_size_stop_cm = pixel_size_stop;
    desc_.los_angle = los_angle;
}

std::vector<std::vector<double>> ImagePlane::getPixelSizes() {
    std::vector<double> flat((size_t)(desc_.n_across * desc_.n_along));
    int rc = jsx_image_pixel_sizes(&desc_, flat.data());
    if (rc) throw Error(rc, "bad image description");
    std::vector<std::vector<double>> out(image_size.first, std::vector<double>(image_size.second));
    for (size_t i = 0; i < image_size.first; ++i)
        for (size_t j = 0; j < image_size.second; ++j) out[i][j] = flat[i * image_size.second + j];
    return out;
}

/* Image::getImage (src/Image.cpp:16-28); pixels never set read as 0 */
std::vector<std::vector<double>> ImagePlane::getImage(std::string value) {
    std::vector<std::vector<double>> out(image_size.first, std::vector<double>(image_size.second, 0.0));
    auto it = planes.find(value);
    if (it == planes.end()) return out;
    for (size_t i = 0; i < image_size.first; ++i)
        for (size_t j = 0; j < image_size.second; ++j) out[i][j] = it->second[i * image_size.second + j];
    return out;
}

/* Image::getStripe (src/Image.cpp:31-41) */
std::vector<double> ImagePlane::getStripe(std::string value) {
    std::vector<double> stripe;
    auto it = planes.find(value);
    for (size_t i = 0; i < image_size.first; ++i)
        for (size_t j = image_size.second / 2; j < image_size.second; ++j)
            if (i == image_size.first / 2)
                stripe.push_back(it == planes.end() ? 0.0 : it->second[i * image_size.second + j]);
    return stripe;
}

/* ---- Observation ---- */
Observation::Observation(Jet *newjet, ImagePlane *newimagePlane, double newnu)
    : nu(newnu), jet(newjet), imagePlane(newimagePlane) {
    std::memset(&params, 0, sizeof params);
    params.stokes = JSX_STOKES_I;
    params.precision = JSX_PREC_FP64;
    params.max_failed_tries = 500;
    params.tau_abs_tol = 1e-7;       /* Observation.cpp:156 */
    params.tau_rel_tol = 1e-14;
    params.stokes_abs_tol = 1e-9;    /* Observation.cpp:269 */
    params.stokes_rel_tol = 1e-9;
    params.tau_precision = 1.0;      /* Observation.cpp:132 */
    params.tau_n_min = 0.1;          /* steps_schedule defaults (utils.h:209-210) */
    params.tau_n_max = 100.0;
    params.n_max_factor = 10;
}

void Observation::run(int n, double tau_max, double dt_max, double tau_min, std::string integration_type,
                      std::string output_type, int n_max, double tau_n_min, double tau_n_max) {
    (void)n_max; (void)tau_n_min; (void)tau_n_max;
    if (integration_type != "adaptive")
        throw Unsupported("integration_type '" + integration_type + "' (only the adaptive path is live)");
    if (output_type != "I" && output_type != "full") throw Unsupported("output_type '" + output_type + "'");
    Context &C = context ? *context : Context::instance();
    jsx_solver_params p = params;
    p.n = n;
    p.tau_max = tau_max;
    p.dt_max = dt_max * pc;           /* Observation.cpp:23 */
    p.tau_min = tau_min;
    p.nu = nu;
    p.stokes = output_type == "full" ? JSX_STOKES_FULL : JSX_STOKES_I;
    jsx_jet_desc jd = jet->descriptor();
    jsx_image_desc id = imagePlane->descriptor();
    size_t N = (size_t)(id.n_across * id.n_along);
    auto &P = imagePlane->planes;
    const char *names[] = {"tau", "I", "Q", "U", "V", "l"};
    for (auto nm : names) P[nm].assign(N, 0.0);
    imagePlane->flags.assign(N, 0u);
    jsx_outputs o;
    std::memset(&o, 0, sizeof o);
    o.tau = P["tau"].data(); o.I = P["I"].data(); o.Q = P["Q"].data(); o.U = P["U"].data(); o.V = P["V"].data();
    o.l = P["l"].data();
    o.flags = imagePlane->flags.data();
    C.check(jsx_observe_host(C.get(), &jd, &id, &p, &o));
    if (output_type == "I") { P.erase("Q"); P.erase("U"); P.erase("V"); }
}

void Observation::run_stripe(int n, double tau_max, double tau_min) {
    (void)tau_min;
    Context &C = context ? *context : Context::instance();
    jsx_solver_params p = params;
    p.n = n;
    p.tau_max = tau_max;
    p.nu = nu;
    jsx_jet_desc jd = jet->descriptor();
    jsx_image_desc id = imagePlane->descriptor();
    std::vector<double> stripe((size_t)(id.n_along - id.n_along / 2), 0.0);
    C.check(jsx_tau_stripe(C.get(), &jd, &id, &p, stripe.data()));
    /* the stripe lands in the "tau" plane's middle row, where getStripe reads */
    auto &tau = imagePlane->planes["tau"];
    tau.assign((size_t)(id.n_across * id.n_along), 0.0);
    const size_t row = (size_t)(id.n_across / 2);
    for (size_t k = 0; k < stripe.size(); ++k) tau[row * id.n_along + id.n_along / 2 + k] = stripe[k];
}

std::vector<std::vector<double>> Observation::getImage(std::string value) { return imagePlane->getImage(value); }
std::vector<double> Observation::getStripe(std::string value) { return imagePlane->getStripe(value); }
std::pair<unsigned long int, unsigned long int> Observation::getImageSize() { return imagePlane->image_size; }

double mas_to_pc(double z) { return jsx_mas_to_pc(z); }

std::ostream &write_2dvector(std::ostream &os, std::vector<std::vector<double>> &v, double scale) {
    for (auto &row : v) {
        for (double x : row) os << x / scale << " ";
        os << "\n";
    }
    return os;
}

std::ostream &write_vector(std::ostream &os, std::vector<double> &v, double scale) {
    for (double x : v) os << x / scale << " " << "\n";
    return os;
}

void read_from_txt(std::string fntxt, std::vector<std::vector<double>> &properties) {
    int64_t n = 0;
    int rc = jsx_read_field_text(fntxt.c_str(), nullptr, 0, &n);
    if (rc == JSX_ENOTFOUND) return;   /* an unopened ifstream reads nothing */
    if (rc) throw Error(rc, "cannot read " + fntxt);
    std::vector<double> buf((size_t)(3 * n));
    if (n && (rc = jsx_read_field_text(fntxt.c_str(), buf.data(), n, &n))) throw Error(rc, "cannot read " + fntxt);
    for (int64_t i = 0; i < n; ++i) properties.push_back({buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]});
}

int32_t create_grid(const std::string &fn, Context &ctx, double nan_value) {
    int64_t n = 0;
    int rc = jsx_read_field_text(fn.c_str(), nullptr, 0, &n);
    if (rc) throw Error(rc, "cannot read " + fn);
    std::vector<double> pts((size_t)(3 * n));
    if (n && (rc = jsx_read_field_text(fn.c_str(), pts.data(), n, &n))) throw Error(rc, "cannot read " + fn);
    pts.resize((size_t)(3 * n));
    return ctx.upload_points(pts, nan_value);
}

}  // namespace jetshow
