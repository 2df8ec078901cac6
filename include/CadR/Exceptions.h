// CadR error classes: the same names and bases as the reference
// (src/CadR/Exceptions.h:13-40).  The C ABI maps them to cadrResult codes.
#pragma once

#include "Export.h"

#include <stdexcept>
#include <string>

namespace CadR {

class CADR_EXPORT Error {
public:
	virtual ~Error() = default;
	virtual const char* what() const noexcept = 0;
};

class CADR_EXPORT LogicError : public Error, public std::logic_error {
public:
	const char* what() const noexcept override { return std::logic_error::what(); }
	explicit LogicError(const char* w) : std::logic_error(w) {}
	explicit LogicError(const std::string& w) : std::logic_error(w) {}
};

class CADR_EXPORT OutOfResources : public Error, public std::runtime_error {
public:
	const char* what() const noexcept override { return std::runtime_error::what(); }
	explicit OutOfResources(const char* w) : std::runtime_error(w) {}
	explicit OutOfResources(const std::string& w) : std::runtime_error(w) {}
};

class CADR_EXPORT Timeout : public Error, public std::runtime_error {
public:
	const char* what() const noexcept override { return std::runtime_error::what(); }
	explicit Timeout(const char* w) : std::runtime_error(w) {}
	explicit Timeout(const std::string& w) : std::runtime_error(w) {}
};

// Thrown for a failing HIP runtime call (the reference's vk::SystemError).
class CADR_EXPORT DeviceError : public Error, public std::runtime_error {
public:
	const char* what() const noexcept override { return std::runtime_error::what(); }
	explicit DeviceError(const std::string& w) : std::runtime_error(w) {}
};

// Throws the exception matching a cadrResult code (no-op for CADR_SUCCESS).
CADR_EXPORT void throwOnError(int cadrResultCode, const char* what);

}  // namespace CadR
