// SDF sampling and its gradient over the built surface (see sesdf_c.h).
#include "handle.h"
