// Core host types of the operator API (mirrors the reference's vedas.h,
// src/vedas.h:41-110, without Thrust/ModernGPU: host vectors are
// std::vector, device data lives behind the C ABI in vedas_hip.h).
#ifndef VEDAS_HOST_VEDAS_H
#define VEDAS_HOST_VEDAS_H

#include <cstdint>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "vedas_hip.h"

#define SCHEMA_SUBJECT 'S'
#define SCHEMA_PREDICATE 'P'
#define SCHEMA_OBJECT 'O'

#define LITERAL_START_ID 5E7
#define EIF_ALPHA 100

typedef uint32_t TYPEID;
typedef unsigned long long DTYPEID;
typedef int TERM_TYPE;
typedef std::unordered_map<std::string, TYPEID> DICTTYPE;
typedef std::unordered_map<TYPEID, std::string> REVERSE_DICTTYPE;

typedef std::vector<TYPEID> TYPEID_HOST_VEC;
typedef std::vector<DTYPEID> DTYPEID_HOST_VEC;
typedef TYPEID_HOST_VEC::iterator TYPEID_HOST_VEC_IT;

typedef std::tuple<size_t, size_t, size_t, size_t> VAR_BOUND;
inline size_t VAR_LB(std::tuple<size_t, size_t> b) { return std::get<0>(b); }
inline size_t VAR_UB(std::tuple<size_t, size_t> b) { return std::get<1>(b); }
inline size_t VAR_LB(VAR_BOUND b) { return std::get<0>(b); }
inline size_t VAR_UB(VAR_BOUND b) { return std::get<1>(b); }
inline size_t VAR_LB2(VAR_BOUND b) { return std::get<2>(b); }
inline size_t VAR_UB2(VAR_BOUND b) { return std::get<3>(b); }

typedef std::map<std::string, std::pair<TYPEID, TYPEID>> VariableBoundMap;

// Error thrown by the operator API when a C-ABI call fails.
class VedasError : public std::runtime_error {
  public:
    VedasError(int code, const std::string &msg)
        : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }
  private:
    int code_;
};

enum ExecuteLogRecordOp { JOIN_OP, UPLOAD_OP, SWAP_OP };

struct ExecuteLogRecord {
    int deviceId;
    ExecuteLogRecordOp op;
    size_t param1, param2, param3;
    std::string paramstr;
    ExecuteLogRecord(int deviceId, ExecuteLogRecordOp op, std::string paramstr,
                     size_t param1, size_t param2 = 0, size_t param3 = 0)
        : deviceId(deviceId), op(op), param1(param1), param2(param2),
          param3(param3), paramstr(std::move(paramstr)) {}
};

#endif
