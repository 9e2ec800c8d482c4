// mock LAMMPS header: see lammps_api.h
#pragma once
#include "lammps_api.h"
