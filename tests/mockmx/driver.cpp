// Calls a mesolf MEX gateway through the mock MEX API.
//   driver <out.bin|-> <arg>...      each arg: "s:d1xd2x...:file" (single from
//   raw file), "d:d1x...:file" (double), "S:value" (double scalar)
// Prints one JSON line: {"status":"ok","dims":[...],"class":...} or
// {"status":"error","id":"...","msg":"..."}; writes the output's raw data to
// out.bin unless "-".
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mex.h"
#include GATEWAY_FILE

static std::vector<mwSize> parse_dims(const std::string& s) {
  std::vector<mwSize> d;
  std::stringstream ss(s);
  std::string t;
  while (std::getline(ss, t, 'x')) d.push_back((mwSize)std::stoull(t));
  return d;
}

static std::string json_escape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    if (c == '\n') { o += "\\n"; continue; }
    o += c;
  }
  return o;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::vector<mxArray*> in;
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("S:", 0) == 0) {
      mwSize one[2] = {1, 1};
      mxArray* m = mxCreateNumericArray(2, one, mxDOUBLE_CLASS, mxREAL);
      *static_cast<double*>(mxGetData(m)) = std::atof(a.c_str() + 2);
      in.push_back(m);
      continue;
    }
    const size_t p1 = a.find(':'), p2 = a.find(':', p1 + 1);
    const mxClassID cls = a[0] == 's' ? mxSINGLE_CLASS : mxDOUBLE_CLASS;
    std::vector<mwSize> d = parse_dims(a.substr(p1 + 1, p2 - p1 - 1));
    mxArray* m = mxCreateNumericArray(d.size(), d.data(), cls, mxREAL);
    m->dims = d;  // keep the dimensions exactly as given
    std::ifstream f(a.substr(p2 + 1), std::ios::binary);
    f.read(reinterpret_cast<char*>(mxGetData(m)), (std::streamsize)m->data.size());
    in.push_back(m);
  }
  mxArray* out[1] = {nullptr};
  try {
    mexFunction(1, out, (int)in.size(), const_cast<const mxArray**>(in.data()));
  } catch (const MexError& e) {
    printf("{\"status\":\"error\",\"id\":\"%s\",\"msg\":\"%s\"}\n", json_escape(e.id).c_str(),
           json_escape(e.what()).c_str());
    return 0;
  }
  printf("{\"status\":\"ok\",\"class\":%d,\"dims\":[", (int)out[0]->cls);
  for (size_t k = 0; k < out[0]->dims.size(); ++k) printf("%s%zu", k ? "," : "", out[0]->dims[k]);
  printf("]}\n");
  if (std::string(argv[1]) != "-") {
    std::ofstream f(argv[1], std::ios::binary);
    f.write(reinterpret_cast<const char*>(out[0]->data.data()), (std::streamsize)out[0]->data.size());
  }
  for (mxArray* m : in) mxDestroyArray(m);
  mxDestroyArray(out[0]);
  return 0;
}
