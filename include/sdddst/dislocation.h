// Value types of the simulator (reference: include/dislocation.h:27-31,
// include/point_defect.h:27-30).
#ifndef SDDDST_CORE_DISLOCATION_H
#define SDDDST_CORE_DISLOCATION_H

namespace sdddstCore {

struct Dislocation {
    double x;
    double y;
    double b;
    bool operator==(Dislocation a) { return a.x == x && a.y == y && a.b == b; }
};

struct PointDefect {
    double x;
    double y;
    bool operator==(PointDefect a) { return a.x == x && a.y == y; }
};

}  // namespace sdddstCore

#endif
