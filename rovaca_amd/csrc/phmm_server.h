#pragma once

#include <cstdint>

#include "../../include/rovaca_phmm.h"

namespace rvc {

int server_start();
void server_stop();
int region_sync_impl(const rvc_region* region, double* out, uint8_t* flags);

}  // namespace rvc
