// concatenate: the output-file merge tool with the reference's command line
// (tools/concatenate.cpp:48-80): concatenate [-N <Nsamp>] infile1 infile2 ... outfile
#include "gpvmc_host.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::printf("Usage:\nconcatenate [-N <Nsamp>] infile1 infile2 ... outfile\n");
        return 1;
    }
    int nsamp = 0;
    std::vector<std::string> in;
    std::string out;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-N") == 0) {
            if (argc < 5 || i + 1 >= argc) {
                std::fprintf(stderr, "Not enough input arguments\n");
                return 1;
            }
            try {
                nsamp = std::stoi(argv[++i]);
            } catch (const std::exception &e) {
                std::fprintf(stderr, "Invalid argument on \"-N\": %s\n", e.what());
                return 1;
            }
        } else if (i == argc - 1) {
            out = argv[i];
        } else {
            in.push_back(argv[i]);
        }
    }
    try {
        gpvmc::Concatenate(in, out, nsamp);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "concatenate: %s\n", e.what());
        return 2;
    }
    return 0;
}
