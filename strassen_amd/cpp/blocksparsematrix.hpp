// blocksparsematrix.hpp -- the drop-in BlockSparseMatrix is header-complete in
// blocksparsematrix.h.
#include "blocksparsematrix.h"
