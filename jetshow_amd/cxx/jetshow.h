/*
 * jetshow.h -- the reference's C++ class API for the per-pixel path, backed by
 * the MI355X C ABI (include/jetshow_amd.h).
 *
 * Field values at single points (bf, nf, vf, is_within, Jet::getKI / getEtaI
 * / getV) are evaluated by the kernels' own field code through
 * jsx_fields_host, one launch per call: they are for inspection, not for
 * loops over many points (Context::fields takes arrays).
 *
 * Class names, constructor arguments and method names follow the reference
 * (include/Geometry.h, Cone.h, Parabaloid.h, Cylinder.h, BField.h, NField.h,
 * VField.h, Jet.h, ImagePlane.h, Observation.h, Ray.h, Intersection.h) so code
 * written against jetshow compiles against this header with the include
 * changed.  Differences:
 *   - Vector3d is a small value type here (Eigen is not a dependency);
 *   - every class can describe itself as a POD descriptor (lower()); kinds the
 *     kernels do not implement throw jetshow::Unsupported instead of running
 *     anywhere else;
 *   - errors from the GPU surface as jetshow::Error exceptions carrying the
 *     jsx_status code (the reference has no error channel).
 */
#ifndef JETSHOW_CXX_H
#define JETSHOW_CXX_H

#include <cstdint>
#include <list>
#include <map>
#include <ostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/jetshow_amd.h"

namespace jetshow {

/* include/utils.h:17-36 */
const double mas_to_rad = 4.84813681109536e-09;
const double rad_to_mas = 206264806.24709633;
const double pc = 3.0856775814671913e+18;
const double m_e = 9.10938356e-28;
const double q_e = 4.80320467299766e-10;
const double c = 29979245800.0;
const double to_jy = 1E+23;
const double pi = 3.141592653589793238462643383279502884;

struct Error : std::runtime_error {
    int code;
    Error(int code_, const std::string &msg) : std::runtime_error(msg), code(code_) {}
};
struct Unsupported : Error {
    explicit Unsupported(const std::string &msg) : Error(JSX_EUNSUPPORTED, msg) {}
};

struct Vector3d {
    double v[3];
    Vector3d() : v{0, 0, 0} {}
    Vector3d(double x, double y, double z) : v{x, y, z} {}
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};

/* ---- Ray / Intersection (src/Ray.cpp, src/Intersection.cpp) ---- */
class Ray {
public:
    Ray(const Vector3d &origin, const Vector3d &direction);  /* normalises direction */
    Vector3d point(double t) const;
    const Vector3d &origin() const { return origin_; }
    const Vector3d &direction() const { return direction_; }

private:
    Vector3d origin_, direction_;
};

class Intersection {
public:
    Intersection(const Vector3d &dir, const Vector3d &pin, const Vector3d &pout)
        : direction_(dir), borders_(pin, pout) {}
    const std::pair<Vector3d, Vector3d> get_path() const { return borders_; }
    const Vector3d direction() const { return direction_; }
    void set_point_out(const Vector3d &p) { borders_.second = p; }
    void set_point_in(const Vector3d &p) { borders_.first = p; }

private:
    Vector3d direction_;
    std::pair<Vector3d, Vector3d> borders_;
};

class Context;

/* ---- geometry (include/Geometry.h:40-55) ---- */
class BaseGeometry {
public:
    virtual ~BaseGeometry() = default;
    /* Geometry::hit on the GPU (jsx_hit_host) */
    std::list<Intersection> hit(Ray &ray) const;
    virtual void lower(jsx_jet_desc &d) const = 0;
};

class Geometry : public BaseGeometry {
public:
    virtual const double big_scale() const = 0;
    virtual const Vector3d &origin() const = 0;
    /* include/Geometry.h:50 (Cone.cpp:48-53, Parabaloid.cpp:35-37,
     * Cylinder.cpp:29-32, ParabaloidCone.cpp:36-43), on the GPU */
    virtual bool is_within(Vector3d &point) const;
};

class Cone : public Geometry {
public:
    Cone(const Vector3d &origin, const Vector3d &direction, const double &angle, const double &scale);
    const Vector3d &origin() const override { return origin_; }
    const Vector3d &direction() const { return direction_; }
    const double angle() const { return angle_; }
    const double big_scale() const override { return big_scale_; }
    void lower(jsx_jet_desc &d) const override;
    /* finite-length extension: keep z in [zmin, zmax] (cm) */
    void set_z_clip(double zmin, double zmax) { zmin_ = zmin; zmax_ = zmax; }

private:
    Vector3d origin_, direction_;
    double angle_, big_scale_, zmin_ = 0, zmax_ = 0;
};

class Parabaloid : public Geometry {
public:
    Parabaloid(const Vector3d &origin, const Vector3d &direction, const double &r0, const double &big_scale);
    const Vector3d &origin() const override { return origin_; }
    const Vector3d &direction() const { return direction_; }
    double r0() const { return r0_; }
    const double big_scale() const override { return big_scale_; }
    void lower(jsx_jet_desc &d) const override;

private:
    Vector3d origin_, direction_;
    double r0_, big_scale_;
};

class Cylinder : public Geometry {
public:
    Cylinder(const Vector3d &origin, const Vector3d &direction, const double &r);
    const Vector3d &origin() const override { return origin_; }
    const Vector3d &direction() const { return direction_; }
    double r() const { return r_; }
    const double big_scale() const override { return 100. * r_; }
    void lower(jsx_jet_desc &d) const override;
    void set_z_clip(double zmin, double zmax) { zmin_ = zmin; zmax_ = zmax; }

private:
    Vector3d origin_, direction_;
    double r_, zmin_ = 0, zmax_ = 0;
};

/* ParabaloidCone (include/ParabaloidCone.h, src/ParabaloidCone.cpp:12-94):
 * the paraboloid r0 below z0, the cone tangent to it there above */
class ParabaloidCone : public Geometry {
public:
    ParabaloidCone(const Vector3d &origin, const Vector3d &direction, const double &r0, const double &z0,
                   const double &scale);
    const Vector3d &origin() const override { return origin_; }
    const Vector3d &direction() const { return direction_; }
    double z0() const { return z0_; }
    const double big_scale() const override { return big_scale_; }
    void lower(jsx_jet_desc &d) const override;

private:
    Vector3d origin_, direction_;
    double r0_, z0_, big_scale_;
};

/* ---- fields (include/BField.h, NField.h, VField.h) ---- */
/* SimulationGeometry (include/Geometry.h:58-66, src/Geometry.cpp:92-132): the
 * convex hull of the simulation points, as face planes uploaded with
 * Context::upload_hull (the hull replaces the reference's CGAL AABB tree) */
class SimulationGeometry : public BaseGeometry {
public:
    explicit SimulationGeometry(int32_t hull_handle) : h_(hull_handle) {}
    void lower(jsx_jet_desc &d) const override;
private:
    int32_t h_;
};

/* include/BField.h:49-53: a tangled field, given in the plasma frame */
class ScalarBField {
public:
    virtual ~ScalarBField() = default;
    virtual void lower(jsx_jet_desc &d) const = 0;
    virtual double bf(const Vector3d &point) const;
    double bf_plasma_frame(const Vector3d &point, Vector3d &v) const { (void)v; return bf(point); }
};
/* include/BField.h:36-44 */
class VectorBField {
public:
    explicit VectorBField(bool in_plasma_frame) : in_plasma_frame_(in_plasma_frame) {}
    virtual ~VectorBField() = default;
    virtual void lower(jsx_jet_desc &d) const = 0;
    virtual Vector3d bf(const Vector3d &point) const;
    /* BField.cpp:212-219: unchanged if given in the plasma frame, else getBprime(b, v) */
    Vector3d bf_plasma_frame(const Vector3d &point, Vector3d &v) const;

protected:
    bool in_plasma_frame_;
};

class RandomScalarBField : public ScalarBField {
public:
    RandomScalarBField(double b_0, double m_b) : b_0_(b_0), m_b_(m_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, m_b_;
};
class RandomScalarBFieldZ : public ScalarBField {
public:
    RandomScalarBFieldZ(double b_0, double m_b) : b_0_(b_0), m_b_(m_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, m_b_;
};
/* CompositeRandomScalarBFieldZ (BField.h:76-86): exponent m_b_inner for
 * |z| <= z0, m_b_outer beyond */
class CompositeRandomScalarBFieldZ : public ScalarBField {
public:
    CompositeRandomScalarBFieldZ(double b_0, double m_b_inner, double m_b_outer, double z0)
        : b_0_(b_0), m_in_(m_b_inner), m_out_(m_b_outer), z0_(z0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, m_in_, m_out_, z0_;
};
class ConstCylinderBField : public VectorBField {
public:
    ConstCylinderBField(double b_0, double n_b, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), n_b_(n_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, n_b_;
};
class RadialConicalBField : public VectorBField {
public:
    RadialConicalBField(double b_0, double n_b, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), n_b_(n_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, n_b_;
};
class HelicalCylinderBField : public VectorBField {
public:
    HelicalCylinderBField(double b_0, double pitch_angle, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), pitch_(pitch_angle) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, pitch_;
};
class SpiralConicalBField : public VectorBField {
public:
    SpiralConicalBField(double b_0, double pitch_angle, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), pitch_(pitch_angle) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, pitch_;
};
class ForceFreeCylindricalBField : public VectorBField {
public:
    ForceFreeCylindricalBField(double b_0, double mu, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), mu_(mu) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, mu_;
};
class ToroidalBField : public VectorBField {
public:
    ToroidalBField(double b_0, double n_b, bool in_plasma_frame) : VectorBField(in_plasma_frame), b_0_(b_0), n_b_(n_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, n_b_;
};
class ConstCylinderBFieldZ : public VectorBField {
public:
    ConstCylinderBFieldZ(double b_0, double n_b, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), n_b_(n_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, n_b_;
};
/* extensions: power law in |p| with a true azimuthal / helical direction */
class ToroidalRBField : public VectorBField {
public:
    ToroidalRBField(double b_0, double n_b, bool in_plasma_frame) : VectorBField(in_plasma_frame), b_0_(b_0), n_b_(n_b) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, n_b_;
};
class HelicalRBField : public VectorBField {
public:
    HelicalRBField(double b_0, double n_b, double pitch, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), b_0_(b_0), n_b_(n_b), pitch_(pitch) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double b_0_, n_b_, pitch_;
};
/* SimulationBField: grid handles from Context::upload_grid */
class SimulationBField : public VectorBField {
public:
    SimulationBField(int32_t grid_p, int32_t grid_fi, bool in_plasma_frame)
        : VectorBField(in_plasma_frame), gp_(grid_p), gf_(grid_fi) {}
    void lower(jsx_jet_desc &d) const override;
private:
    int32_t gp_, gf_;
};
class SimulationScalarBField : public ScalarBField {
public:
    explicit SimulationScalarBField(int32_t grid) : g_(grid) {}
    void lower(jsx_jet_desc &d) const override;
private:
    int32_t g_;
};

/* include/NField.h:10-18 */
class NField {
public:
    explicit NField(bool in_plasma_frame) : in_plasma_frame_(in_plasma_frame) {}
    virtual ~NField() = default;
    virtual void lower(jsx_jet_desc &d) const = 0;
    virtual double nf(const Vector3d &point) const;
    /* NField.cpp:7-14 */
    double nf_plasma_frame(const Vector3d &point, double &gamma) const {
        double n = nf(point);
        return in_plasma_frame_ ? n : n / gamma;
    }
protected:
    bool in_plasma_frame_;
};
class ConstNField : public NField {
public:
    ConstNField(double n, bool in_plasma_frame) : NField(in_plasma_frame), n_(n) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double n_;
};
class BKNField : public NField {
public:
    BKNField(double n_0, double n_n, bool in_plasma_frame) : NField(in_plasma_frame), n_0_(n_0), n_n_(n_n) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double n_0_, n_n_;
};
class BKNFieldZ : public NField {
public:
    BKNFieldZ(double n_0, double n_n, bool in_plasma_frame) : NField(in_plasma_frame), n_0_(n_0), n_n_(n_n) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double n_0_, n_n_;
};
/* CompositeBKNField (NField.h:50-60): exponent n_n_inner for |z| <= z0, n_n_outer beyond */
class CompositeBKNField : public NField {
public:
    CompositeBKNField(double n_0, double n_n_inner, double n_n_outer, double z0, bool in_plasma_frame)
        : NField(in_plasma_frame), n_0_(n_0), n_in_(n_n_inner), n_out_(n_n_outer), z0_(z0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double n_0_, n_in_, n_out_, z0_;
};
class SimulationNField : public NField {
public:
    SimulationNField(int32_t grid, bool in_plasma_frame) : NField(in_plasma_frame), g_(grid) {}
    void lower(jsx_jet_desc &d) const override;
private:
    int32_t g_;
};

/* include/VField.h:11-14 */
class VField {
public:
    virtual ~VField() = default;
    virtual void lower(jsx_jet_desc &d) const = 0;
    virtual Vector3d vf(const Vector3d &point) const;
};
class ConstFlatVField : public VField {
public:
    explicit ConstFlatVField(double gamma) : gamma_(gamma) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double gamma_;
};
class ConstCentralVField : public VField {
public:
    explicit ConstCentralVField(double gamma, Vector3d origin = Vector3d(0, 0, 0)) : gamma_(gamma), origin_(origin) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double gamma_;
    Vector3d origin_;
};
/* include/VField.h:27-176 */
class ShearedFlatVField : public VField {
public:
    ShearedFlatVField(double gamma_axis, double gamma_border, double r) : ga_(gamma_axis), gb_(gamma_border), r_(r) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double ga_, gb_, r_;
};
class SheathFlatVField : public VField {
public:
    SheathFlatVField(double gamma_spine, double gamma_sheath, double r_sheath)
        : ga_(gamma_spine), gb_(gamma_sheath), r_(r_sheath) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double ga_, gb_, r_;
};
/* the reference's theta = acos(z + o_z / r) (VField.cpp:83) is NaN off the
 * origin; the polar angle acos((z + o_z) / r) is used unless the solver
 * flag JSX_OPT_SHEARED_CENTRAL_REF_BUG asks for the reference's form */
class ShearedCentralVField : public VField {
public:
    ShearedCentralVField(double gamma_axis, double gamma_border, double theta, Vector3d origin = Vector3d(0, 0, 0))
        : ga_(gamma_axis), gb_(gamma_border), theta_(theta), origin_(origin) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double ga_, gb_, theta_;
    Vector3d origin_;
};
class SheathCentralVField : public VField {
public:
    SheathCentralVField(double gamma_spine, double gamma_sheath, double theta_sheath,
                        Vector3d origin = Vector3d(0, 0, 0))
        : ga_(gamma_spine), gb_(gamma_sheath), theta_(theta_sheath), origin_(origin) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double ga_, gb_, theta_;
    Vector3d origin_;
};
class ConstParabolicVField : public VField {
public:
    ConstParabolicVField(double gamma, double Rz0) : g_(gamma), Rz0_(Rz0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double g_, Rz0_;
};
class AccParabolicVField : public VField {
public:
    AccParabolicVField(double gamma0, double R0, double Rz0) : g_(gamma0), R0_(R0), Rz0_(Rz0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double g_, R0_, Rz0_;
};
class ShearedAccParabolicVField : public VField {
public:
    ShearedAccParabolicVField(double gamma_axis0, double gamma_border0, double R0, double Rz0)
        : ga_(gamma_axis0), gb_(gamma_border0), R0_(R0), Rz0_(Rz0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double ga_, gb_, R0_, Rz0_;
};
class AccParabolicConstConeVField : public VField {
public:
    AccParabolicConstConeVField(double gamma0, double Rz0, double z0) : g_(gamma0), Rz0_(Rz0), z0_(z0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double g_, Rz0_, z0_;
};
class ShearedAccParabolicConstConeVField : public VField {
public:
    ShearedAccParabolicConstConeVField(double gamma_axis0, double gamma_border0, double Rz0, double z0)
        : ga_(gamma_axis0), gb_(gamma_border0), Rz0_(Rz0), z0_(z0) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double ga_, gb_, Rz0_, z0_;
};
class LinearGammaZVField : public VField {
public:
    LinearGammaZVField(double gamma0, double gamma1, double z0, double z1) : g0_(gamma0), g1_(gamma1), z0_(z0), z1_(z1) {}
    void lower(jsx_jet_desc &d) const override;
private:
    double g0_, g1_, z0_, z1_;
};
class SimulationVField : public VField {
public:
    explicit SimulationVField(int32_t grid) : g_(grid) {}
    void lower(jsx_jet_desc &d) const override;
private:
    int32_t g_;
};

/* ---- Jet (include/Jet.h:11-61) ---- */
class Jet {
public:
    Jet(BaseGeometry *geo, VField *vfield, ScalarBField *bField, NField *nField);
    Jet(BaseGeometry *geo, VField *vfield, VectorBField *bField, NField *nField);
    std::list<Intersection> hit(Ray &ray);
    /* include/Jet.h:21,38,50 (src/Jet.cpp:15-46,218-237,290-296): lab-frame
     * k_I, eta_I and bulk velocity at a point; and the restored getters of
     * the commented full-Stokes code (src/Jet.cpp:114-214,239-288) in the sky
     * Stokes frame */
    double getKI(Vector3d &point, Vector3d &n_los, double nu);
    double getEtaI(Vector3d &point, Vector3d &n_los, double nu);
    const Vector3d getV(const Vector3d &point);
    double getKQ(Vector3d &point, Vector3d &n_los, double nu);
    double getKU(Vector3d &point, Vector3d &n_los, double nu);
    double getKV(Vector3d &point, Vector3d &n_los, double nu);
    double getKF(Vector3d &point, Vector3d &n_los, double nu);
    double getKC(Vector3d &point, Vector3d &n_los, double nu);
    double gethQ(Vector3d &point, Vector3d &n_los, double nu);
    double getEtaQ(Vector3d &point, Vector3d &n_los, double nu);
    double getEtaU(Vector3d &point, Vector3d &n_los, double nu);
    double getEtaV(Vector3d &point, Vector3d &n_los, double nu);
    /* solver option bits used when the getters evaluate (e.g.
     * JSX_OPT_SHEARED_CENTRAL_REF_BUG); Observation takes its own */
    uint32_t flags = 0;
    /* emission parameters the reference hard-codes (s = 3.42, Jet.cpp:32) */
    void set_s(double s) { s_ = s; }
    void set_gamma_min(double g) { gamma_min_ = g; }
    void set_thermal_fraction(double f) { thermal_ = f; }
    jsx_jet_desc descriptor() const;

private:
    BaseGeometry *geometry_;
    VField *vfield_;
    ScalarBField *sbfield_ = nullptr;
    VectorBField *vbfield_ = nullptr;
    NField *nfield_;
    double s_ = 3.42, gamma_min_ = 100.0, thermal_ = 0.0;
};

/* ---- ImagePlane (include/ImagePlane.h:22-35) ---- */
class ImagePlane {
public:
    ImagePlane(std::pair<unsigned long int, unsigned long int> image_size, double pixel_size_start,
               double pixel_size_stop, double los_angle);
    std::vector<std::vector<double>> getImage(std::string value);
    std::vector<std::vector<double>> getPixelSizes();
    std::vector<double> getStripe(std::string value);
    const std::pair<unsigned long int, unsigned long int> image_size;
    jsx_image_desc descriptor() const { return desc_; }
    /* result planes written by Observation::run, row-major [n_across * n_along] */
    std::map<std::string, std::vector<double>> planes;
    std::vector<uint32_t> flags;

private:
    jsx_image_desc desc_;
};

/* ---- GPU context shared by the classes ---- */
class Context {
public:
    explicit Context(int device = 0, size_t hbm_budget = 0);
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    jsx_ctx *get() const { return ctx_; }
    /* jsx_fields_host: JSX_FIELDS_STRIDE values per point */
    std::vector<double> fields(const jsx_jet_desc &d, const std::vector<double> &points,
                               const std::vector<double> &n_los, double nu, uint32_t flags = 0);
    int32_t upload_grid(const std::vector<double> &values, int nz, int nr, double z0, double dz, double r0,
                        double dr, double nan_value);
    /* nodes: n x (z [cm], r [cm], value) at any positions (jsx_grid_upload_points) */
    int32_t upload_points(const std::vector<double> &zrv, double nan_value);
    /* planes: n_planes x (n_x, n_y, n_z, h) in cm, inside where n . p <= h */
    int32_t upload_hull(const std::vector<double> &planes);
    static Context &instance();   /* process-wide default context on device 0 */
    void check(int rc) const;

private:
    jsx_ctx *ctx_ = nullptr;
};

/* ---- Observation (include/Observation.h:12-52) ---- */
class Observation {
public:
    Observation(Jet *newjet, ImagePlane *imagePlane, double nu);
    /* src/Observation.cpp:20-41: dt_max is multiplied by pc here, as the
     * reference does; n_max, tau_n_min, tau_n_max are accepted and, as in the
     * reference's adaptive path, not used. integration_type must be
     * "adaptive"; output_type "I" or "full". */
    void run(int n, double tau_max, double dt_max, double tau_min, std::string integration_type,
             std::string output_type, int n_max, double tau_n_min, double tau_n_max);
    /* src/Observation.cpp:310-347: optical depth of the middle row's jet-side
     * pixels with the fixed-step RK4 integrate_tau (n steps per segment);
     * getStripe("tau") returns it.  tau_min is accepted and unused, as in the
     * reference. */
    void run_stripe(int n, double tau_max, double tau_min);
    std::vector<std::vector<double>> getImage(std::string value);
    std::vector<double> getStripe(std::string value);
    std::pair<unsigned long int, unsigned long int> getImageSize();
    const double nu;
    /* solver parameters beyond the reference signature */
    jsx_solver_params params;
    Context *context = nullptr;

private:
    Jet *jet;
    ImagePlane *imagePlane;
};

/* frame transforms of include/utils.h (src/utils.cpp:147-168) */
double getG(const Vector3d &v);
double getD(const Vector3d &n_los, const Vector3d &v);
Vector3d get_n_los_prime(const Vector3d &n_los, const Vector3d &v);
Vector3d getBprime(const Vector3d &b, const Vector3d &v);

/* mas_to_pc (src/utils.cpp:205-208) */
double mas_to_pc(double z);

/* text formats (include/utils.h:192-198, src/utils.cpp:211-236,381-388):
 * write_2dvector / write_vector print value/scale with the stream's current
 * format and a space, rows (values) end with a newline; read_from_txt reads
 * rows of three numbers until the first that does not parse. */
std::ostream &write_2dvector(std::ostream &os, std::vector<std::vector<double>> &v, double scale = 1.0);
std::ostream &write_vector(std::ostream &os, std::vector<double> &v, double scale = 1.0);
void read_from_txt(std::string fntxt, std::vector<std::vector<double>> &properties);
/* create_triangulation (src/SimulationInterpolater.cpp:105-117) for a text
 * file of field nodes at any positions: the nodes in HBM (a complete regular
 * (z, r) grid in the regular layout, any other set Delaunay-triangulated),
 * their handle for the Simulation*Field classes */
int32_t create_grid(const std::string &fn, Context &ctx, double nan_value = 0.0);

}  // namespace jetshow

#endif
