// blockallocator.hpp -- see blockallocator.h.
#include "blockallocator.h"
