// host_util.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <string>

#include "rtx.h"

namespace rtx {

// records msg as the calling thread's rtx_last_error() and returns st
rtx_status set_error(rtx_status st, const std::string& msg);

}  // namespace rtx
