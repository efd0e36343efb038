// gdlog/exception.h -- error convention of the C++ API.
//
// The reference aborts the process on any CUDA error (checkCuda ->
// cudaDeviceReset -> exit, include/exception.cuh:6-19).  Here every engine
// call returns a gdl_status; the C++ API turns a failure into a gdlog::Error
// exception, which terminates a program that does not catch it (the
// reference's behaviour) and lets callers that do catch it recover.
#pragma once

#include <stdexcept>
#include <string>

#include "gdl_hip.h"

namespace gdlog {

class Error : public std::runtime_error {
  public:
    Error(gdl_status st, const std::string &msg) : std::runtime_error(msg), status(st) {}
    gdl_status status;
};

// Throws gdlog::Error when st != GDL_OK, with the context's last message.
void check(gdl_status st, const char *what);

}  // namespace gdlog
