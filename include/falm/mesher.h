// Falm C++ host API: Mesher (src/mesher/mesher.hpp:9-262), which turns a
// node list (mesh.json + x.txt, y.txt, z.txt) into a control-volume file.
// The conversion is falm_mesher_build of the C ABI (host code).
#ifndef FALM_AMD_MESHER_H
#define FALM_AMD_MESHER_H

#include <string>

#include "typedef.h"

namespace Falm {

class Mesher {
public:
    // cvcenter "innerNode" or "cell"; returns FALM_OK or FALM_ERR_SETUP
    // (message in falm_last_error) where the reference's would throw
    static int build_mesh(const std::string &cvcenter, const std::string &meshpath, const std::string &outputfile,
                          int gc) {
        return falmErrCheckMacro(falm_mesher_build(cvcenter.c_str(), meshpath.c_str(), outputfile.c_str(), gc));
    }
};

}  // namespace Falm

// the reference declares Mesher at global scope
using Mesher = Falm::Mesher;

#endif
