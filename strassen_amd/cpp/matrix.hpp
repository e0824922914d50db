// matrix.hpp -- the drop-in Matrix is header-complete in matrix.h.
#include "matrix.h"
