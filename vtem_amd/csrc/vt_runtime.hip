#include "vt_common.hpp"
