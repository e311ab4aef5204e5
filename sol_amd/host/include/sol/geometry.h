// The intersection boundary, as sol::Geometry (reference include/sol/geometry.h:33-43).
// sol intersects one ray per virtual call on the CPU; here queries are batched and run on the
// GPU through solhip_intersect (see TriangleMesh::intersect_closest / intersect_any).
#ifndef SOL_GEOMETRY_H
#define SOL_GEOMETRY_H

namespace sol {

class Geometry {
public:
    virtual ~Geometry() = default;
};

} // namespace sol

#endif
