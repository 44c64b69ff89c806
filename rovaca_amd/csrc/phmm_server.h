#pragma once

#include <cstdint>

#include "../../include/rovaca_phmm.h"

namespace rvc {

int server_start();
void server_stop();
bool server_running();
// One region through the serving loop (submit + wait), as a worker thread
// would call it.
int served_region(const rvc_region* region, double* out, uint8_t* flags, const DevicePrep* prep = nullptr);
struct DevicePrep;
int region_sync_impl(const rvc_region* region, double* out, uint8_t* flags, const DevicePrep* prep = nullptr);
// The no-allocation one-region call's set-up (its device scratch address and
// stream) at init, and its release at shutdown (phmm_noalloc.cpp).
int noalloc_init();
void noalloc_shutdown();

}  // namespace rvc
