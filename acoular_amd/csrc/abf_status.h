// Error reporting shared by the C-ABI translation units: the message of the last failure on the
// calling thread, returned by abf_last_error().
#pragma once

namespace abf {

// record msg as the calling thread's last error and return code
int set_error(int code, const char *msg);

} // namespace abf
