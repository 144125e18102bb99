// ORACLE — voxel part (filled in below).
#include "fpt_oracle.h"
