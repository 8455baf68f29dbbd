// Deconvolves an unsigned 16-bit raw volume with a 16-bit raw kernel on the
// GPU: the reference's example app (apps/DeconvolveRawImages/
// DeconvolveRawImages.cxx:36-133) over this library's C API.
//
//   DeconvolveRawImages <image> <x> <y> <z> <kernel> <x> <y> <z> <output>
//                       [--iterations N] [--clamp]
//
// Same command line, file format (little-endian uint16, x fastest, no
// header), method (Clarity_MaximumLikelihoodDeconvolve, 5 iterations, the
// kernel's raw counts cast to float) and report lines as the reference.
// The float -> uint16 store keeps the reference's unclamped cast
// (DeconvolveRawImages.cxx:116) with the result x86-64 gives it: truncate
// toward zero to a 32-bit integer (0x80000000 when out of range or NaN) and
// keep the low 16 bits, so values above 65535 wrap.  --clamp saturates to
// [0, 65535] instead.  Unlike the reference, a failed read or a failed
// deconvolution ends the program with status 1 and no output written.
#include <Clarity.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

bool read_u16(const char *path, size_t n, std::vector<float> &out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::vector<uint16_t> raw(n);
  const size_t got = std::fread(raw.data(), sizeof(uint16_t), n, f);
  std::fclose(f);
  if (got != n) return false;
  out.resize(n);
  for (size_t i = 0; i < n; i++) {
    const unsigned char *b = reinterpret_cast<const unsigned char *>(&raw[i]);
    out[i] = static_cast<float>(static_cast<uint16_t>(b[0] | (b[1] << 8)));
  }
  return true;
}

// x86-64 static_cast<unsigned short>(float): cvttss2si to int32, low 16 bits
uint16_t cast_like_reference(float v) {
  int32_t i;
  if (std::isnan(v) || v >= 2147483648.0f || v < -2147483648.0f)
    i = INT32_MIN;
  else
    i = static_cast<int32_t>(v);  // truncation toward zero
  return static_cast<uint16_t>(static_cast<uint32_t>(i) & 0xffffu);
}

uint16_t cast_clamped(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 65535.0f) return 65535;
  return static_cast<uint16_t>(v);
}

}  // namespace

int main(int argc, char *argv[]) {
  if (argc < 10) {
    std::printf(
        "Usage: DeconvolveRawImages <image filename> <x> <y> <z> <psf filename> <x> <y> <z> "
        "<output file name> [--iterations N] [--clamp]\n");
    return 1;
  }
  const char *image_name = argv[1];
  Clarity_Dim3 image_dim{std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
  const char *psf_name = argv[5];
  Clarity_Dim3 psf_dim{std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8])};
  const char *output_name = argv[9];
  int iterations = 5;
  bool clamp = false;
  for (int a = 10; a < argc; a++) {
    if (!std::strcmp(argv[a], "--clamp")) {
      clamp = true;
    } else if (!std::strcmp(argv[a], "--iterations") && a + 1 < argc) {
      iterations = std::atoi(argv[++a]);
    } else {
      std::printf("ERROR: unknown option '%s'\n", argv[a]);
      return 1;
    }
  }
  if (image_dim.x <= 0 || image_dim.y <= 0 || image_dim.z <= 0 || psf_dim.x <= 0 ||
      psf_dim.y <= 0 || psf_dim.z <= 0 || iterations < 0) {
    std::printf("ERROR: dimensions must be positive and iterations non-negative\n");
    return 1;
  }
  const size_t n_image = (size_t)image_dim.x * image_dim.y * image_dim.z;
  const size_t n_psf = (size_t)psf_dim.x * psf_dim.y * psf_dim.z;

  std::vector<float> image, psf;
  if (!read_u16(image_name, n_image, image)) {
    std::printf("ERROR: Could not read %zu 16-bit values from image file '%s'\n", n_image,
                image_name);
    return 1;
  }
  if (!read_u16(psf_name, n_psf, psf)) {
    std::printf("ERROR: Could not read %zu 16-bit values from PSF file '%s'\n", n_psf,
                psf_name);
    return 1;
  }

  std::vector<float> result(n_image);
  const ClarityResult_t rc = Clarity_MaximumLikelihoodDeconvolve(
      image.data(), image_dim, psf.data(), psf_dim, result.data(), iterations);
  if (rc != CLARITY_SUCCESS) {
    std::printf("ERROR: deconvolution failed: %s\n", Clarity_ResultString(rc));
    return 1;
  }

  std::vector<unsigned char> out(2 * n_image);
  for (size_t i = 0; i < n_image; i++) {
    const uint16_t v = clamp ? cast_clamped(result[i]) : cast_like_reference(result[i]);
    out[2 * i] = static_cast<unsigned char>(v & 0xff);
    out[2 * i + 1] = static_cast<unsigned char>(v >> 8);
  }
  FILE *f = std::fopen(output_name, "wb");
  if (!f || std::fwrite(out.data(), 1, out.size(), f) != out.size()) {
    std::printf("ERROR: Could not write file '%s'\n", output_name);
    if (f) std::fclose(f);
    return 1;
  }
  std::fclose(f);

  std::printf("Input image file name: %s\n", image_name);
  std::printf("Input image dimensions: %d x %d x %d\n", image_dim.x, image_dim.y, image_dim.z);
  std::printf("Input point-spread function file name: %s\n", psf_name);
  std::printf("Input point-spread function dimensions: %d x %d x %d\n", psf_dim.x, psf_dim.y,
              psf_dim.z);
  std::printf("Output file name: %s\n", output_name);
  return 0;
}
