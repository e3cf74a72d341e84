// growth.h — Smoluchowski coagulation kernels (declared here, defined in growth.hip).
#pragma once
struct shark_growth;
void shark_growth_free(shark_growth* g);
