/*
 * The class API used the way the reference's drivers use it: fields and
 * geometries constructed with the reference's constructor arguments, their
 * virtual bf / nf / vf / is_within, Jet's getters, and an
 * Observation::run on a cone model as src/main.cpp sets one up
 * (Cone, RandomScalarBFieldZ, BKNFieldZ, ConstCentralVField through a VField
 * pointer, ImagePlane of log-spaced pixels, getPixelSizes, getImage).
 * Prints "tag value..." lines (%.17g) that tests/test_gpu_class_api.py
 * compares with the oracle.
 */
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "jetshow.h"

using namespace jetshow;

static void put(const char *tag, const Vector3d &v) { std::printf("%s %.17g %.17g %.17g\n", tag, v[0], v[1], v[2]); }
static void put(const char *tag, double x) { std::printf("%s %.17g\n", tag, x); }

int main() {
    Vector3d origin = {0., 0., 0.};
    Vector3d direction = {0., 0., 1.};
    Vector3d p(0.02 * pc, 0.01 * pc, 2.0 * pc);
    Vector3d q(-0.05 * pc, 0.03 * pc, 4.0 * pc);

    /* B fields: include/BField.h */
    RandomScalarBField b1(0.1, 1.0);
    RandomScalarBFieldZ b2(0.1, 1.0);
    CompositeRandomScalarBFieldZ b3(0.1, 1.0, 2.0, 3.0 * pc);
    ConstCylinderBField b4(0.1, 1.0, true);
    ConstCylinderBFieldZ b5(0.1, 1.0, true);
    RadialConicalBField b6(0.1, 1.0, false);
    ToroidalBField b7(0.1, 1.0, true);
    HelicalCylinderBField b8(0.05, 0.5, true);
    SpiralConicalBField b9(0.5, 0.3, true);
    ForceFreeCylindricalBField b10(0.05, 1.5, true);
    put("RandomScalarBField", b1.bf(p));
    put("RandomScalarBFieldZ", b2.bf(p));
    put("CompositeRandomScalarBFieldZ.p", b3.bf(p));
    put("CompositeRandomScalarBFieldZ.q", b3.bf(q));
    put("ConstCylinderBField", b4.bf(p));
    put("ConstCylinderBFieldZ", b5.bf(p));
    put("RadialConicalBField", b6.bf(p));
    put("ToroidalBField", b7.bf(p));
    put("HelicalCylinderBField", b8.bf(p));
    put("SpiralConicalBField", b9.bf(p));
    put("ForceFreeCylindricalBField", b10.bf(p));
    Vector3d v_lab(0.0, 0.0, 0.9 * c);
    put("RadialConicalBField.plasma", b6.bf_plasma_frame(p, v_lab));
    put("ToroidalBField.plasma", b7.bf_plasma_frame(p, v_lab));
    put("RandomScalarBField.plasma", b1.bf_plasma_frame(p, v_lab));

    /* N fields: include/NField.h */
    ConstNField n1(300.0, true);
    BKNField n2(500.0, 2.0, true);
    BKNFieldZ n3(500.0, 2.0, false);
    CompositeBKNField n4(500.0, 2.0, 3.0, 3.0 * pc, true);
    put("ConstNField", n1.nf(p));
    put("BKNField", n2.nf(p));
    put("BKNFieldZ", n3.nf(p));
    put("CompositeBKNField.p", n4.nf(p));
    put("CompositeBKNField.q", n4.nf(q));
    double gam = 5.0;
    put("BKNFieldZ.plasma", n3.nf_plasma_frame(p, gam));

    /* V fields: include/VField.h */
    ConstFlatVField v1(5.0);
    ShearedFlatVField v2(10.0, 3.0, 1.0 * pc);
    SheathFlatVField v3(10.0, 3.0, 0.015 * pc);
    ConstCentralVField v4(5.0);
    ShearedCentralVField v5(10.0, 3.0, 0.05);
    SheathCentralVField v6(10.0, 3.0, 0.015);
    ConstParabolicVField v7(5.0, 0.1 * pc);
    AccParabolicVField v8(10.0, 5.0 * pc, 0.1 * pc);
    ShearedAccParabolicVField v9(10.0, 3.0, 5.0 * pc, 0.1 * pc);
    AccParabolicConstConeVField v10(10.0, 0.1 * pc, 2.0 * pc);
    ShearedAccParabolicConstConeVField v11(10.0, 3.0, 0.1 * pc, 2.0 * pc);
    put("ConstFlatVField", v1.vf(p));
    put("ShearedFlatVField", v2.vf(p));
    put("SheathFlatVField", v3.vf(p));
    put("ConstCentralVField", v4.vf(p));
    put("ShearedCentralVField", v5.vf(p));
    put("SheathCentralVField", v6.vf(p));
    put("ConstParabolicVField", v7.vf(p));
    put("AccParabolicVField", v8.vf(p));
    put("ShearedAccParabolicVField", v9.vf(p));
    put("AccParabolicConstConeVField.p", v10.vf(p));
    put("AccParabolicConstConeVField.q", v10.vf(q));
    put("ShearedAccParabolicConstConeVField.q", v11.vf(q));

    /* geometries: is_within and hit */
    Cone cone(origin, direction, 0.03, 100 * pc);
    Parabaloid para(origin, direction, 0.1 * pc, 100 * pc);
    Cylinder cyl(origin, direction, 0.3 * pc);
    ParabaloidCone pcone(origin, direction, 0.1 * pc, 2.0 * pc, 100 * pc);
    put("Cone.is_within", cone.is_within(p) ? 1.0 : 0.0);
    put("Cylinder.is_within", cyl.is_within(q) ? 1.0 : 0.0);
    put("Parabaloid.is_within", para.is_within(p) ? 1.0 : 0.0);
    put("ParabaloidCone.is_within.p", pcone.is_within(p) ? 1.0 : 0.0);
    put("ParabaloidCone.is_within.q", pcone.is_within(q) ? 1.0 : 0.0);
    Ray ray(Vector3d(0.01 * pc, 0.0, 20.0 * pc), Vector3d(-std::sin(0.3), 0.0, -std::cos(0.3)));
    auto segs = pcone.hit(ray);
    put("ParabaloidCone.hit.n", (double)segs.size());
    if (!segs.empty()) {
        put("ParabaloidCone.hit.in", segs.front().get_path().first);
        put("ParabaloidCone.hit.out", segs.front().get_path().second);
    }

    /* Jet getters on a helical-B paraboloid-cone jet */
    HelicalRBField hb(0.1, 1.0, 0.7853981633974483, false);
    BKNField nb(500.0, 2.0, true);
    Jet jet(&pcone, &v10, &hb, &nb);
    jet.set_s(2.5);
    Vector3d n_los(std::sin(0.1), 0.0, std::cos(0.1));
    double nu = 15.4e9;
    put("Jet.getV", jet.getV(p));
    put("Jet.getKI", jet.getKI(p, n_los, nu));
    put("Jet.getEtaI", jet.getEtaI(p, n_los, nu));
    put("Jet.getKQ", jet.getKQ(p, n_los, nu));
    put("Jet.getKF", jet.getKF(p, n_los, nu));
    put("Jet.getEtaV", jet.getEtaV(p, n_los, nu));

    /* frame transforms of utils.h */
    Vector3d vv = jet.getV(p);
    put("getG", getG(vv));
    put("getD", getD(n_los, vv));
    put("get_n_los_prime", get_n_los_prime(n_los, vv));
    put("getBprime", getBprime(hb.bf(p), vv));

    /* an Observation as src/main.cpp sets one up */
    double los_angle = 0.3, redshift = 0.05;
    Cone geometry(origin, direction, 0.05, 100);
    RandomScalarBFieldZ bfield(0.2, 1.0);
    BKNFieldZ nfield(2e3, 2.0, true);
    VField *vfield = new ConstCentralVField(8.0);
    Jet bkjet(&geometry, vfield, &bfield, &nfield);
    auto image_size = std::make_pair((unsigned long)16, (unsigned long)24);
    double pc_in_mas = mas_to_pc(redshift);
    double lg_start = std::log10(0.005 * pc_in_mas * pc), lg_stop = std::log10(0.05 * pc_in_mas * pc);
    ImagePlane imagePlane(image_size, lg_start, lg_stop, los_angle);
    auto sizes = imagePlane.getPixelSizes();
    put("pixel_size.0.0", sizes[0][0]);
    put("pixel_size.last", sizes[15][23]);
    double nu_observed_ghz = 8.1;
    nu_observed_ghz *= 1E+09;
    nu_observed_ghz *= (1. + redshift);
    Observation observation(&bkjet, &imagePlane, nu_observed_ghz);
    std::string step_type = "adaptive", calculate = "I";
    observation.run(100, 1e4, pc * 0.001, std::pow(10., -10.0), step_type, calculate, 2000, 0.1, 1e4);
    auto tau = observation.getImage("tau");
    auto I = observation.getImage("I");
    double tsum = 0, isum = 0, imax = 0;
    for (auto &row : tau) for (double x : row) tsum += x;
    for (auto &row : I) for (double x : row) { isum += x; imax = x > imax ? x : imax; }
    put("Observation.tau.sum", tsum);
    put("Observation.I.sum", isum);
    put("Observation.I.max", imax);
    delete vfield;
    return 0;
}
