// growth.h — dust-growth state owned by a step context (defined in growth.hip).
#pragma once
struct shark_growth;
void shark_growth_free(shark_growth* g);
namespace shark {
int fail_msg(int code, const char* msg);
}
