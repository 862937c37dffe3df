/*
 * pyjetshow -- the reference's Python binding (src/pyjetshow.cpp:69-201),
 * same function name, arguments, defaults and return value, running through
 * the class API on the GPU.  get_i_image_np returns numpy arrays instead of
 * nested lists (plus the per-pixel flags) for callers that fit many models.
 */
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <fstream>
#include <vector>

#include "jetshow.h"

namespace py = pybind11;
using namespace jetshow;

struct IImage {
    std::vector<std::vector<double>> tau, I;
    std::vector<uint32_t> flags;
};

/* src/pyjetshow.cpp:69-173 */
static IImage run_i_image(double los_angle, double redshift, unsigned long number_of_pixels_along,
                          unsigned long number_of_pixels_across, double pixel_size_mas_start,
                          double pixel_size_mas_stop, double cone_half_angle, double B_1, double m, double K_1,
                          double n, double Gamma, double nu_observed_ghz, double tau_max, bool central_vfield) {
    Vector3d origin(0., 0., 0.), direction(0., 0., 1.);
    Cone geometry(origin, direction, cone_half_angle, 100);   /* big_scale as passed at :79 */
    RandomScalarBFieldZ bfield(B_1, m);
    BKNFieldZ nfield(K_1, n, true);
    ConstCentralVField vcentral(Gamma);
    ConstFlatVField vflat(Gamma);
    VField *vfield = central_vfield ? (VField *)&vcentral : (VField *)&vflat;
    Jet bkjet(&geometry, vfield, &bfield, &nfield);
    auto image_size = std::make_pair(number_of_pixels_across, number_of_pixels_along);
    double pc_in_mas = mas_to_pc(redshift);
    double lg_pixel_size_start = std::log10(pixel_size_mas_start * pc_in_mas * pc);
    double lg_pixel_size_stop = std::log10(pixel_size_mas_stop * pc_in_mas * pc);
    ImagePlane imagePlane(image_size, lg_pixel_size_start, lg_pixel_size_stop, los_angle);
    auto pixel_sizes = imagePlane.getPixelSizes();
    double nu = nu_observed_ghz * 1E+09;
    nu *= (1. + redshift);
    Observation observation(&bkjet, &imagePlane, nu);
    double dt_max = pc * 0.001;
    double tau_min = std::pow(10., -10.0);
    observation.run(100, tau_max, dt_max, tau_min, "adaptive", "I", 2000, 0.1, tau_max);
    IImage r;
    r.tau = observation.getImage("tau");
    r.I = observation.getImage("I");
    r.flags = imagePlane.flags;
    for (size_t i = 0; i < r.I.size(); ++i)
        for (size_t j = 0; j < r.I[i].size(); ++j) {
            /* src/pyjetshow.cpp:111-134,163-169 */
            double omega = (pixel_sizes[i][j] / (pc_in_mas * pc)) * (pixel_sizes[i][j] / (pc_in_mas * pc)) *
                           mas_to_rad * mas_to_rad;
            double scale = 1E-23 * (1. + redshift) * (1. + redshift) * (1. + redshift) / omega;
            r.I[i][j] = r.I[i][j] / scale;
        }
    return r;
}

static std::pair<std::vector<std::vector<double>>, std::vector<std::vector<double>>>
get_i_image(double los_angle, double redshift, unsigned long number_of_pixels_along,
            unsigned long number_of_pixels_across, double pixel_size_mas_start, double pixel_size_mas_stop,
            double cone_half_angle, double B_1, double m, double K_1, double n, double Gamma,
            double nu_observed_ghz, double tau_max, bool central_vfield) {
    IImage r = run_i_image(los_angle, redshift, number_of_pixels_along, number_of_pixels_across, pixel_size_mas_start,
                           pixel_size_mas_stop, cone_half_angle, B_1, m, K_1, n, Gamma, nu_observed_ghz, tau_max,
                           central_vfield);
    return std::make_pair(r.tau, r.I);
}

static py::tuple get_i_image_np(double los_angle, double redshift, unsigned long number_of_pixels_along,
                                unsigned long number_of_pixels_across, double pixel_size_mas_start,
                                double pixel_size_mas_stop, double cone_half_angle, double B_1, double m,
                                double K_1, double n, double Gamma, double nu_observed_ghz, double tau_max,
                                bool central_vfield) {
    IImage r = run_i_image(los_angle, redshift, number_of_pixels_along, number_of_pixels_across, pixel_size_mas_start,
                           pixel_size_mas_stop, cone_half_angle, B_1, m, K_1, n, Gamma, nu_observed_ghz, tau_max,
                           central_vfield);
    size_t na = number_of_pixels_across, nl = number_of_pixels_along;
    py::array_t<double> tau({na, nl}), I({na, nl});
    py::array_t<uint32_t> fl({na, nl});
    auto t = tau.mutable_unchecked<2>();
    auto ii = I.mutable_unchecked<2>();
    auto f = fl.mutable_unchecked<2>();
    for (size_t i = 0; i < na; ++i)
        for (size_t j = 0; j < nl; ++j) {
            t(i, j) = r.tau[i][j];
            ii(i, j) = r.I[i][j];
            f(i, j) = r.flags[i * nl + j];
        }
    return py::make_tuple(tau, I, fl);
}

PYBIND11_MODULE(pyjetshow, m) {
    using namespace pybind11::literals;
    m.doc() = "Radiative transfer for BK models (MI355X)";
    auto args = [](auto &&f, const char *name, const char *doc, py::module_ &mod) {
        mod.def(name, f, doc, "los_angle"_a, "redshift"_a, "number_of_pixels_along"_a, "number_of_pixels_across"_a,
                "pixel_size_mas_start"_a, "pixel_size_mas_stop"_a, "cone_half_angle"_a, "B_1"_a, "m"_a, "K_1"_a,
                "n"_a, "Gamma"_a, "nu_observed_ghz"_a, "tau_max"_a = 10000., "central_vfield"_a = false);
    };
    args(&get_i_image, "get_i_image", "Obtain Stokes I image with random B-field", m);
    args(&get_i_image_np, "get_i_image_np", "get_i_image returning numpy arrays (tau, I_Jy, flags)", m);
    m.def("mas_to_pc", &mas_to_pc, "parsec per milliarcsecond at redshift z", "z"_a);
    /* the text formats through the class API's stream writers */
    m.def(
        "write_2dvector",
        [](const std::string &path, std::vector<std::vector<double>> v, double scale, bool append) {
            std::fstream fs;
            fs.open(path, std::ios::out | (append ? std::ios::app : std::ios::trunc));
            if (!fs.is_open()) throw std::runtime_error("cannot open " + path);
            write_2dvector(fs, v, scale);
        },
        "write_2dvector into a file (src/utils.cpp:211-224, src/main.cpp:1447-1462)", "path"_a, "v"_a,
        "scale"_a = 1.0, "append"_a = true);
    m.def(
        "write_vector",
        [](const std::string &path, std::vector<double> v, double scale, bool append) {
            std::fstream fs;
            fs.open(path, std::ios::out | (append ? std::ios::app : std::ios::trunc));
            if (!fs.is_open()) throw std::runtime_error("cannot open " + path);
            write_vector(fs, v, scale);
        },
        "write_vector into a file (src/utils.cpp:227-236)", "path"_a, "v"_a, "scale"_a = 1.0, "append"_a = true);
    m.def(
        "read_from_txt",
        [](const std::string &path) {
            std::vector<std::vector<double>> rows;
            read_from_txt(path, rows);
            return rows;
        },
        "rows of three numbers (src/utils.cpp:381-388)", "path"_a);
}
