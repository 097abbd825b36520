// json_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own JSON (de)serializer (G4HepEm/G4HepEmDataJsonIO,
// entry points G4HepEmDataJsonIO.cc:45-61), compiled from the reference tree
// by oracle/Makefile into oracle/_ref/libhepem_json_ref.so, on this project's
// tables. It pins the library's JSON reader and writer (hepem_json.cpp) in
// both directions: documents the reference writes must load there, and
// documents written there must load in the reference.
//
// The project's table structs are reinterpreted as the reference's; their
// member offsets are checked in ref_harness.cc against the same headers.

#include "G4HepEmData.hh"
#include "G4HepEmDataJsonIO.hh"
#include "G4HepEmParameters.hh"
#include "G4HepEmState.hh"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>

extern "C" {

// G4HepEmStateToJson of (data, pars); *out is malloc'd text of *len bytes.
int refjson_write_state(const void* data, const void* pars, char** out, uint64_t* len) {
  G4HepEmState st;
  st.fData = const_cast<G4HepEmData*>(static_cast<const G4HepEmData*>(data));
  st.fParameters = const_cast<G4HepEmParameters*>(static_cast<const G4HepEmParameters*>(pars));
  std::ostringstream os;
  if (!G4HepEmStateToJson(os, &st)) return -1;
  const std::string s = os.str();
  *out = static_cast<char*>(std::malloc(s.size() ? s.size() : 1));
  if (!*out) return -1;
  std::memcpy(*out, s.data(), s.size());
  *len = s.size();
  return 0;
}

void refjson_free_text(char* p) { std::free(p); }

// G4HepEmStateFromJson; NULL if the reference rejects the text.
void* refjson_read_state(const char* text, uint64_t len) {
  try {
    std::istringstream is(std::string(text, len));
    return G4HepEmStateFromJson(is);
  } catch (...) {
    return nullptr;
  }
}
void* refjson_state_data(void* st) { return static_cast<G4HepEmState*>(st)->fData; }
void* refjson_state_pars(void* st) { return static_cast<G4HepEmState*>(st)->fParameters; }
void refjson_free_state(void* p) {
  G4HepEmState* st = static_cast<G4HepEmState*>(p);
  if (!st) return;
  if (st->fData) {
    FreeG4HepEmData(st->fData);
    delete st->fData;
  }
  if (st->fParameters) {
    FreeG4HepEmParameters(st->fParameters);
    delete st->fParameters;
  }
  delete st;
}

}  // extern "C"
