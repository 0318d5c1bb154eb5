// Errors of the host library, mirroring src/error.rs of the reference:
// ProgramError kinds (error.rs:30-46), VegasError::NotImplementedError
// (error.rs:25-26), IoError (error.rs:48-56), plus the device failures of the
// GPU backend.  Raised as
// exceptions (the C++ counterpart of the reference's Result<_, Error>).
#pragma once

#include <stdexcept>
#include <string>

#include "vegas_hip.h"

namespace vegas {

enum class ErrorKind {
    TemperatureMaxLessThanMin = VG_ERR_TEMPERATURE_MAX_LESS_THAN_MIN,
    NoSteps = VG_ERR_NO_STEPS,
    ZeroTemperature = VG_ERR_ZERO_TEMPERATURE,
    ZeroCoolRate = VG_ERR_ZERO_COOL_RATE,
    ZeroField = VG_ERR_ZERO_FIELD,
    ZeroFieldStep = VG_ERR_ZERO_FIELD_STEP,
    NotImplemented = VG_ERR_NOT_IMPLEMENTED,
    InvalidArgument = VG_ERR_INVALID_ARGUMENT,
    UnsupportedLattice = VG_ERR_UNSUPPORTED_LATTICE,
    Device = VG_ERR_DEVICE,
    OutOfMemory = VG_ERR_OUT_OF_MEMORY,
    NoDevice = VG_ERR_NO_DEVICE,
    // Host-only: files of the output tables (IoError, error.rs:48-56).  The
    // C ABI does no file I/O, so no status code maps here.
    Io = 1000,
};

// Messages of error.rs:32-43.
inline const char* error_message(ErrorKind k) {
    switch (k) {
        case ErrorKind::TemperatureMaxLessThanMin:
            return "maximum temperature must be greater than minimum temperature";
        case ErrorKind::NoSteps: return "there should be at least one step";
        case ErrorKind::ZeroTemperature: return "temperature must be greater than zero";
        case ErrorKind::ZeroCoolRate: return "cooling rate must be greater than zero";
        case ErrorKind::ZeroField: return "maximum field must be greater than zero";
        case ErrorKind::ZeroFieldStep: return "field step must be greater than zero";
        case ErrorKind::NotImplemented: return "not implemented error";
        case ErrorKind::Io: return "io error";
        default: return "device error";
    }
}

class Error : public std::runtime_error {
  public:
    Error(ErrorKind kind, const std::string& what) : std::runtime_error(what), kind_(kind) {}
    explicit Error(ErrorKind kind) : Error(kind, error_message(kind)) {}
    ErrorKind kind() const { return kind_; }

  private:
    ErrorKind kind_;
};

// Throws the library's last error for a non-zero status of the C ABI.
inline void check(int status) {
    if (status != VG_OK) throw Error(static_cast<ErrorKind>(status), vg_last_error());
}

}  // namespace vegas
