// Interior and boundary mode lists of the static condensation (schur.hpp).
#include "schur.hpp"

namespace nektar_amd {
namespace schur {

namespace {
// flags over the coefficient order: hex (p, q, r) with p fastest, tet
// (p, q, r) with r fastest and q < nm - p, r < nm - p - q
std::vector<char> interior_flags(Shape shape, int nm)
{
    std::vector<char> f;
    if (shape == Shape::Hex) {
        for (int r = 0; r < nm; ++r)
            for (int q = 0; q < nm; ++q)
                for (int p = 0; p < nm; ++p) f.push_back(p >= 2 && q >= 2 && r >= 2);
    } else {
        for (int p = 0; p < nm; ++p)
            for (int q = 0; q < nm - p; ++q)
                for (int r = 0; r < nm - p - q; ++r) f.push_back(p >= 2 && q >= 1 && r >= 1);
    }
    return f;
}

std::vector<int> select(Shape shape, int nm, bool interior)
{
    const std::vector<char> f = interior_flags(shape, nm);
    std::vector<int> out;
    for (int i = 0; i < (int)f.size(); ++i)
        if ((f[i] != 0) == interior) out.push_back(i);
    return out;
}
}  // namespace

std::vector<int> interior_modes(Shape shape, int nm) { return select(shape, nm, true); }
std::vector<int> boundary_modes(Shape shape, int nm) { return select(shape, nm, false); }

}  // namespace schur
}  // namespace nektar_amd
