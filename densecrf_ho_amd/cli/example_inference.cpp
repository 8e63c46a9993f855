// example_inference on the GPU engine: the reference's command line
// (src/example_inference.cpp:15-134) over the C++ interface.
//
//   example_inference method [spc_std spc_potts bil_spcstd bil_colstd bil_potts [const_1 norm_1]]
//                     [--image PATH] [--unary PATH] [--output PATH] [--regions PATH]
//
// method: unary, qp, lp, qp_sp, lp_sp.  Defaults as the reference: image
// ../data/img.bmp, unaries ../data/img.c_unary, output ../data/seg.png,
// spatial std 2 and Potts 3, bilateral spatial std 30, colour std 8, Potts 2,
// superpixel constant 50 and normaliser 1000.  The superpixel methods take
// the segmentation as a region map (--regions: one little-endian int32
// region id per pixel, row-major), which the reference computes with
// mean-shift (spatial radius 8, range radius 4, minimum region 100).
// Prints "# method: <method>\t<discretised energy>" as the reference does.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "densecrf_ho_amd_io.hpp"

using namespace dcrf_ho;

static std::vector<int> read_regions(const std::string& path, size_t n, int* R) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("cannot open " + path);
  std::vector<int> r(n);
  const size_t got = std::fread(r.data(), sizeof(int), n, fp);
  std::fclose(fp);
  if (got != n) throw std::runtime_error("region map: expected one int32 per pixel in " + path);
  int mx = -1;
  for (int v : r) {
    if (v < 0) throw std::runtime_error("region map: negative region id");
    mx = v > mx ? v : mx;
  }
  *R = mx + 1;
  return r;
}

int main(int argc, char* argv[]) {
  std::vector<std::string> pos;
  std::string image = "../data/img.bmp", unaries = "../data/img.c_unary", output = "../data/seg.png", regions;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error(a + " needs a value");
      return argv[++i];
    };
    if (a == "--image") image = next();
    else if (a == "--unary") unaries = next();
    else if (a == "--output") output = next();
    else if (a == "--regions") regions = next();
    else pos.push_back(a);
  }
  if (pos.empty()) {
    std::cout << "Usage: ./example_inference method spc_std spc_potts bil_spcstd bil_colstd bil_potts [const_1 norm_1]"
                 " [--image P] [--unary P] [--output P] [--regions P]\n";
    return 1;
  }
  try {
    const std::string method = pos[0];
    float spc_std = 2.0f, spc_potts = 3.0f, bil_potts = 2.0f, bil_spcstd = 30.0f, bil_colstd = 8.0f;
    if (pos.size() == 6 || pos.size() == 8) {
      spc_std = std::stof(pos[1]);
      spc_potts = std::stof(pos[2]);
      bil_spcstd = std::stof(pos[3]);
      bil_colstd = std::stof(pos[4]);
      bil_potts = std::stof(pos[5]);
    }
    float const_1 = 50.f, norm_1 = 1000.f;
    if (pos.size() == 8) {
      const_1 = std::stof(pos[6]);
      norm_1 = std::stof(pos[7]);
    }
    LP_inf_params lp_params;
    lp_params.prox_energy_tol = lp_params.dual_gap_tol;

    img_size size = {-1, -1};
    MatrixXf unary = load_unary(unaries, size);
    std::unique_ptr<unsigned char[]> img(load_image(image, size));
    DenseCRF2D crf(size.width, size.height, (int)unary.rows());
    crf.setUnaryEnergy(unary);
    crf.addPairwiseGaussian(spc_std, spc_std, new PottsCompatibility(spc_potts));
    crf.addPairwiseBilateral(bil_spcstd, bil_spcstd, bil_colstd, bil_colstd, bil_colstd, img.get(),
                             new PottsCompatibility(bil_potts));
    MatrixXf Q = crf.unary_init();
    const bool sp = method == "qp_sp" || method == "lp_sp";
    if (sp) {
      if (regions.empty())
        throw std::runtime_error(method + " needs the superpixel segmentation as a region map (--regions)");
      int R = 0;
      std::vector<int> reg = read_regions(regions, (size_t)size.width * size.height, &R);
      crf.addSuperPixel(img.get(), reg, R, const_1, norm_1);
    }
    if (method == "lp_sp") {
      std::cout << "---Running tests on proximal LP with higher order potentials\r\n";
      try {
        Q = crf.lp_inference_prox_super_pixels(Q, lp_params);
      } catch (std::runtime_error& e) {
        std::cout << "Runtime Error!: " << e.what() << std::endl;
      }
    } else if (method == "lp") {
      std::cout << "---Running tests on proximal LP\r\n";
      Q = crf.lp_inference_prox_super_pixels(Q, lp_params);
    } else if (method == "qp") {
      std::cout << "---Running tests on QP with non convex energy\r\n";
      Q = crf.qp_inference_super_pixels_non_convex(Q);
    } else if (method == "qp_sp") {
      std::cout << "---Running tests on QP with non convex energy with higher order potentials\r\n";
      Q = crf.qp_inference_super_pixels_non_convex(Q);
    } else if (method != "unary") {
      std::cout << "Unrecognised method.\n";
    }
    // the reference takes the higher-order energy for qp_sp only (its
    // "prox_lp_sp" branch names no method)
    const double e = method == "qp_sp" ? crf.assignment_energy_higher_order(crf.currentMap(Q))
                                       : crf.assignment_energy_true(crf.currentMap(Q));
    save_map(Q, size, output, "MSRC");
    std::printf("# method: %s\t%.17g\n", method.c_str(), e);
  } catch (std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
