// Intermediate result base class (reference IR.h:5-10).
#ifndef VEDAS_HOST_IR_H
#define VEDAS_HOST_IR_H

class IR {
  public:
    IR() = default;
    virtual ~IR() = default;
};

#endif
