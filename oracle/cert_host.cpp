// Host build of the device builder's certified atan2 and sine (test
// infrastructure only): dioptrin_amd/csrc/atan2_cert.hpp and sin_cert.hpp
// compiled for the CPU with -ffp-contract=off, beside glibc's atan2 and sin,
// so tests can check that every result a certificate accepts equals libm bit
// for bit.
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../dioptrin_amd/csrc/atan2_cert.hpp"
#include "../dioptrin_amd/csrc/sin_cert.hpp"

extern "C" {

// For each (y[k], x[k]): cert[k] = atan2_cert's value, libm[k] = glibc's
// atan2, flag[k] = 1 when the certificate rejects its value (the device
// builder then takes the host libm's).
void or_atan2_cert(const double* y, const double* x, size_t n, double* cert, double* libm,
                   uint8_t* flag) {
  for (size_t k = 0; k < n; k++) {
    const dio::Atan2Result r = dio::atan2_cert(y[k], x[k]);
    cert[k] = r.value;
    flag[k] = r.uncertain ? 1 : 0;
    libm[k] = std::atan2(y[k], x[k]);
  }
}

// For each x[k]: cert[k] = sin_cert's value, libm[k] = glibc's sin, flag[k]
// = 1 when the certificate rejects its value.
void or_sin_cert(const double* x, size_t n, double* cert, double* libm, uint8_t* flag) {
  for (size_t k = 0; k < n; k++) {
    const dio::SinResult r = dio::sin_cert(x[k]);
    cert[k] = r.value;
    flag[k] = r.uncertain ? 1 : 0;
    libm[k] = std::sin(x[k]);
  }
}
}
