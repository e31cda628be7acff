// PointDefect lives with Dislocation (reference: include/point_defect.h).
#ifndef SDDDST_CORE_POINT_DEFECT_H
#define SDDDST_CORE_POINT_DEFECT_H
#include "dislocation.h"
#endif
