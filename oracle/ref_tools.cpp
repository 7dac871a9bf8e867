// Test infrastructure: a command-line driver for the parts of the reference
// that compile on their own, built into oracle/_ref/ by oracle/build.py from
// the reference's sources where they lie (it needs the reference's include
// directory, src/, on the include path):
//   - Mesher::build_mesh (src/mesher/mesher.hpp:12-260), header-only;
//   - its vendored JSON library (src/nlohmann/json.hpp, nlohmann/json
//     3.11.3), which writes the reference's setup copies, index files and
//     actuator-point files (src/FalmIO.h:19-67, src/alm/bladeHandler.h:47-103).
//
//   ref_tools mesh CVCENTER MESHPATH OUTFILE GC
//   ref_tools json FILE        parse, print dump(2) (sorted keys)
//   ref_tools ojson FILE       parse as ordered_json, print dump(2)
//   ref_tools num V...         print json(double V).dump() per line
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "mesher/mesher.hpp"

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "mesh" && argc == 6) {
        Mesher::build_mesh(argv[2], argv[3], argv[4], std::atoi(argv[5]));
        return 0;
    }
    if ((mode == "json" || mode == "ojson") && argc == 3) {
        std::ifstream f(argv[2]);
        if (mode == "json") std::cout << nlohmann::json::parse(f).dump(2);
        else std::cout << nlohmann::ordered_json::parse(f).dump(2);
        return 0;
    }
    if (mode == "num") {
        for (int a = 2; a < argc; a++) std::cout << nlohmann::json(std::strtod(argv[a], nullptr)).dump() << "\n";
        return 0;
    }
    return 2;
}
