// Symbol visibility of the CadR facade classes in libcadr_hip.so (the library
// is built with hidden visibility; only the C ABI and these classes export).
#pragma once

#ifndef CADR_EXPORT
#define CADR_EXPORT __attribute__((visibility("default")))
#endif
