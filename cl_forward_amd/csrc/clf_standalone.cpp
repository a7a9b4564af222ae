// clf_standalone.cpp — the reference's standalone program
// (IndependentEntrypoint.cpp:145-193) over this library: reads framed .dat
// files from a folder, round robin up to the requested count, runs
// independent_execute (x-presort in place, then the search) and
// independent_post_execute.
//
//   clf_standalone <folder with .dat files> [number of files=50]
//                  [--results DIR] [--sorted]
//
// --results DIR writes DIR/<i>.out per event in the reference's printTrack
// format (what its PRINT_VERBOSE build writes to results/, KernelInvoker.cpp
// :209-243).  --sorted takes the files in name order; by default they come
// in directory order, as the reference's readdir loop takes them.
#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "GpuPixelSearchByTriplet.h"
#include "clf.h"

namespace {

void print_usage(char* argv[]) {
    std::cerr << "Usage: " << argv[0]
              << " <folder including .dat files> [number of files to process=50]"
                 " [--results DIR] [--sorted]"
              << std::endl;
}

// IndependentEntrypoint.cpp:89-123: int funcNameLen; funcName; int dataSize; data
bool read_framed(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path.c_str(), std::ifstream::binary);
    if (!f) {
        std::cerr << "Error: File " << path << " does not exist." << std::endl;
        return false;
    }
    std::vector<uint8_t> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const uint8_t* data = nullptr;
    size_t n = 0;
    if (clf_unframe(raw.data(), raw.size(), nullptr, nullptr, &data, &n) != CLF_OK) {
        std::cerr << "Error: " << path << ": " << clf_last_error() << std::endl;
        return false;
    }
    out.assign(data, data + n);
    return true;
}

}  // namespace

int main(int argc, char* argv[]) {
    std::vector<std::string> pos;
    std::string results_dir;
    bool sorted = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--results" && i + 1 < argc) {
            results_dir = argv[++i];
        } else if (a == "--sorted") {
            sorted = true;
        } else {
            pos.push_back(a);
        }
    }
    if (pos.empty()) {
        print_usage(argv);
        return -1;
    }
    const std::string folder = pos[0];
    size_t n_files = 50;
    if (pos.size() >= 2) n_files = (size_t)std::atoi(pos[1].c_str());

    std::cout << "Execution options:" << std::endl
              << " Executable: " << argv[0] << std::endl
              << " Folder name: " << folder << std::endl
              << " Number of files: " << n_files << std::endl
              << std::endl;

    std::vector<std::string> names;
    if (DIR* dir = opendir(folder.c_str())) {
        while (struct dirent* ent = readdir(dir)) {
            const std::string fn = ent->d_name;
            const size_t dot = fn.find_last_of('.');
            if (dot != std::string::npos && fn.substr(dot + 1) == "dat") names.push_back(fn);
        }
        closedir(dir);
    }
    if (names.empty()) {
        std::cerr << "Directory " << folder << " either could not be opened, or is empty"
                  << std::endl;
        return -1;
    }
    if (sorted) std::sort(names.begin(), names.end());

    std::vector<std::vector<uint8_t> > input(n_files);
    for (size_t i = 0; i < n_files; ++i)
        if (!read_framed(folder + "/" + names[i % names.size()], input[i])) return -1;
    std::cout << input.size() << " files read" << std::endl;

    std::vector<std::vector<uint8_t> > output;
    independent_execute(input, output);
    independent_post_execute(output);

    if (!results_dir.empty()) {
        for (size_t i = 0; i < output.size(); ++i) {
            size_t len = 0;
            const int64_t nt = (int64_t)(output[i].size() / sizeof(clf_track));
            clf_status s = clf_format_tracks(input[i].data(), input[i].size(), output[i].data(), nt,
                                             nullptr, 0, &len);
            std::string text(len + 1, '\0');
            if (s == CLF_OK)
                s = clf_format_tracks(input[i].data(), input[i].size(), output[i].data(), nt,
                                      &text[0], text.size(), &len);
            if (s != CLF_OK) {
                std::cerr << "event " << i << ": " << clf_last_error() << std::endl;
                return -1;
            }
            std::ofstream out(results_dir + "/" + std::to_string(i) + ".out");
            out.write(text.data(), (std::streamsize)len);
        }
    }
    return 0;
}
