// Command-line search over the C++ host API, mirroring the reference's main
// (src/search.cpp:9-26): read the parameters, load the SBWT into HBM, search
// every k-mer of every line of the query file, print the answers.
#include "../../include/sbwt_gpu_amd.hpp"

int main(int argc, char** argv) {
  sbwt_amd::Parameters parameters;
  if (parameters.read_arguments(argc, argv)) return 1;
  sbwt_amd::DeviceMemory device_memory;
  sbwt_amd::MatrixSBWT sbwt;
  if (sbwt.create(device_memory, parameters)) return 1;
  sbwt_amd::KmerSearch kmer_search(parameters.fileQueries, parameters);
  if (kmer_search.search(device_memory)) return 1;
  return 0;
}
