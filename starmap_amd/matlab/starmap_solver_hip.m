function output = starmap_solver_hip(par, precision)
%STARMAP_SOLVER_HIP   STARMAP_SOLVER with its time steps on an AMD MI355X.
%   output = starmap_solver_hip(par) is a drop-in for starmap_solver(par)
%   (starmap_solver.m:1): the same problem struct from starmap_init, the
%   same output callbacks (which may return par), summary print and struct
%   array of final moments.  The loop control stays here, as in the
%   reference -- flags, grids, time step, material/source/filter handles
%   evaluated on the staggered grids -- and the Strang substeps run on the
%   GPU through the MEX gateway starmap_hip_mex (libstarmap_hip).
%   output = starmap_solver_hip(par, 'fp32') keeps the state in FP32.
%
%   Not exercised in this repository (no MATLAB in its image); the Python
%   mirror starmap_amd/solver.py runs the same sequence of library calls
%   and is tested against the oracle.

if nargin < 2, precision = 'fp64'; end
n_sys = length(par.Mz);

% Flags (starmap_solver.m:58-71)
f_order = [];
flag_filter = isfield(par,'filterfunction') && isfield(par,'f_position') ...
    && isfield(par,'mom_order');
if flag_filter
    f_order = par.mom_order;
else
    par.filterfunction = @ones; par.f_position = '';
end
flag_ani = isfield(par,'sigma_sm') && isfield(par,'mom_order');
if ~flag_ani
    par.mom_order = [0 ones(1,n_sys-1)]; par.sigma_sm = @zero;
end
if strcmp(par.mom_output,'all'), par.mom_output = 1:n_sys; end
if ~isfield(par,'timestep'), par.timestep = 0; end

% Device handle: grid types, storage, schedules (:76-98, :136-146)
hh = (par.ax([2 4 6]) - par.ax([1 3 5])) ./ par.n;
H = starmap_hip_mex('create', par.Mx, par.My, par.Mz, par.n, par.bc, hh, ...
    par.mom_order, f_order, par.f_position, precision, 0);
release = onCleanup(@() starmap_hip_mex('destroy', H));
gt = starmap_hip_mex('grid_types', H);
even = find(mod(sum(gt-1,2),2)==0)'; odd = find(mod(sum(gt-1,2),2)==1)';

% Time step bound and parabolic symbol (:101-110)
dt_hyp = min(hh)/abs(eigs(par.Mz,1,'lm'));
if par.n(1)~=1, dt_hyp = dt_hyp/3; elseif par.n(2)~=1, dt_hyp = dt_hyp/2; end
M = -1i*(par.Mx/hh(1)+par.My/hh(2)+par.Mz/hh(3));
time_dependent = [nargin(par.sigma_a)>3, nargin(par.sigma_s0)>3, nargin(par.sigma_sm)>4];
td_source = ~isempty(par.source_ind) && nargin(par.source)>3;
td_filter = flag_filter && nargin(par.filterfunction)>6;

% Staggered grids (:113-126)
ax = {par.ax(1:2), par.ax(3:4), par.ax(5:6)};
g = cell(3,2);
for d = 1:3
    g{d,1} = ax{d}(1)+hh(d)/2:hh(d):ax{d}(2)-hh(d)/2;
    g{d,2} = ax{d}(1)+hh(d)*(1-par.bc(d)):hh(d):ax{d}(2);
end
X = cell(2,2,2); Y = X; Z = X;
for a = 1:2, for b = 1:2, for c = 1:2
    [X{a,b,c},Y{a,b,c},Z{a,b,c}] = ndgrid(g{1,a},g{2,b},g{3,c});
end, end, end
sg = unique([gt(2:end,:) par.mom_order(2:end)'],'rows');     % (grid, order)
f_sg = [];
if flag_filter, f_sg = unique([gt(2:end,:) f_order(2:end)'],'rows'); end

% Initial state (:140-146)
for j = 1:n_sys
    if nargin(par.ic)>3 || j==1
        G = num2cell(gt(j,:));
        starmap_hip_mex('set_state', H, j, X{G{:}}*0 + ...
            capargs(par.ic,X{G{:}},Y{G{:}},Z{G{:}},j));
    end
end

% Time loop (:162-475)
par.t_plot = [par.t_plot inf]; plot_count = 1;
cputime = zeros(1,3); t = 0; dt = 0; step_count = 0; pending = 0;
sA = []; sS = containers.Map(); sT = [];
while t < par.tfinal
    tic
    tleft = (par.tfinal-t)/par.cn;
    if t==0 || any(time_dependent)
        pending = flush(H, dt, pending);
        % materials and time step (:168-266)
        for k = double(t>0):1
            check = k>0 && par.timestep>0;
            if check
                dt = max(dt,dt_hyp);
                decay = decay_rates(sA, sT, gt, par.mom_order);
                [dt_b, la0] = stable_dt(par, M, even, odd, decay, dt, tleft);
            else
                dt_b = dt_hyp;
            end
            if k>0
                dt = min(dt_b,tleft)*par.cn;
                sA0 = sA; sT0 = sT;
                if ~any(time_dependent), sA1 = sA{1,1,1}; sT1 = sT; break, end
            end
            ev = time_dependent | (dt==0 & [1 1 flag_ani]);
            [sA, sS, sT] = materials(par, X, Y, Z, sg, t+dt, ev, sA, sS, sT);
            if k>0
                sA1 = (sA0{1,1,1}+sA{1,1,1})/2;
                sT1 = sT;
                for q = 1:size(sg,1), sT1{q} = (sT0{q}+sT{q})/2; end
            end
            if check && ~(growth_factor(par,M,even,odd, ...
                    decay_rates({sA1}, sT1, gt, par.mom_order),dt) < la0)
                sA1 = sA0{1,1,1}; sT1 = sT0;
            end
        end
        starmap_hip_mex('set_coeff', H, 'sigma_a1', [1 1 1], 0, 0, sA1);
        for q = 1:size(sg,1)
            starmap_hip_mex('set_coeff', H, 'sigma_t', sg(q,1:3), sg(q,4), 0, sT1{q});
        end
    end
    t = t+dt; step_count = step_count+1;
    if step_count==1 || td_source                                  % :276-282
        pending = flush(H, dt, pending);
        for j = par.source_ind
            G = num2cell(gt(j,:));
            starmap_hip_mex('set_coeff', H, 'source', [1 1 1], j, 0, ...
                X{G{:}}*0 + capargs(par.source,X{G{:}},Y{G{:}},Z{G{:}},t-dt/2,j));
        end
    end
    if flag_filter && (t==dt || td_filter)                          % :284-305
        pending = flush(H, dt, pending);
        for q = 1:size(f_sg,1)
            G = num2cell(f_sg(q,1:3));
            fk = @(tt) capargs(par.filterfunction,par,dt,f_sg(q,4),X{G{:}},Y{G{:}},Z{G{:}},tt);
            if strcmp(par.f_position,'full')
                starmap_hip_mex('set_coeff', H, 'filter', f_sg(q,1:3), f_sg(q,4), 0, fk(t));
            else
                starmap_hip_mex('set_coeff', H, 'filter', f_sg(q,1:3), f_sg(q,4), 1, fk(t-dt/2));
                starmap_hip_mex('set_coeff', H, 'filter', f_sg(q,1:3), f_sg(q,4), 2, fk(t));
            end
        end
    end
    cputime(2) = cputime(2)+toc;
    if t >= par.t_plot(plot_count)-1e-14                             % :423-474
        pending = flush(H, dt, pending);
        starmap_hip_mex('snapshot', H, par.mom_output);
        tic, starmap_hip_mex('step', H, dt, 1); cputime(1) = cputime(1)+toc;
        tic
        while t >= par.t_plot(plot_count)-1e-14
            lambda = (par.t_plot(plot_count)-t+dt)/dt;
            Uplot = starmap_hip_mex('interp', H, lambda, par.mom_output);
            par = call_output(par, Uplot, gt, g, plot_count);
            plot_count = plot_count+1;
        end
        cputime(3) = cputime(3)+toc;
    else
        pending = pending+1;
    end
end
tic, pending = flush(H, dt, pending); cputime(1) = cputime(1)+toc; %#ok<NASGU>
starmap_hip_mex('check_finite', H);

% Summary (:477-482)
fprintf('%s with %s%s%d\n %d moments, %dx%dx%d grid, %0.0f time steps\n',...
    par.name,par.closure_mod,par.closure,par.n_mom,n_sys,par.n,step_count)
cputime = reshape([cputime;cputime/sum(cputime)*1e2],1,[]);
fprintf(['CPU-times\n advection:%12.2fs%5.0f%%\n ',...
    'absorb/scatter:%7.2fs%5.0f%%\n plotting:%13.2fs%5.0f%%\n'],cputime)

% Final state (:484-500)
if nargout
    U = cell(1,n_sys);
    for j = 1:n_sys, U{j} = starmap_hip_mex('get_state', H, j); end
    xs = g(1,gt(:,1)); ys = g(2,gt(:,2)); zs = g(3,gt(:,3));
    switch par.dim
        case 1
            for j = 1:n_sys, U{j} = reshape(U{j},length(zs{j}),[]); end
            output = struct('x',zs,'U',U);
        case 2
            for j = 1:n_sys, U{j} = reshape(U{j},length(ys{j}),[]); end
            output = struct('x',ys,'y',zs,'U',U);
        otherwise
            output = struct('x',xs,'y',ys,'z',zs,'U',U);
    end
end
end

%========================================================================
function pending = flush(H, dt, pending)
% Steps held back while no host work was due run as one library call.
if pending > 0, starmap_hip_mex('step', H, dt, pending); end
pending = 0;
end

function [sA, sS, sT] = materials(par, X, Y, Z, sg, tt, ev, sA, sS, sT)
% sigma_a on every grid, sigma_s0 - sigma_sm and sigma_t on the (grid,
% order) pairs sg, re-evaluated where ev says so (:214-243); sS keeps its
% value between calls, as in the reference.
if ev(1), sA = cellfun(@(x,y,z)capargs(par.sigma_a,x,y,z,tt),X,Y,Z,'Un',0); end
for q = 1:size(sg,1)
    G = num2cell(sg(q,1:3)); key = sprintf('%d', q);
    if ev(2), sS(key) = capargs(par.sigma_s0,X{G{:}},Y{G{:}},Z{G{:}},tt); end
    if ev(3)
        sS(key) = sS(key) - capargs(par.sigma_sm,X{G{:}},Y{G{:}},Z{G{:}},sg(q,4),tt);
    end
end
if any(ev)
    sT = cell(1,size(sg,1));
    for q = 1:size(sg,1)
        G = num2cell(sg(q,1:3));
        sT{q} = sA{G{:}} + sS(sprintf('%d', q));
    end
end
end

function decay = decay_rates(sA, sT, gt, mom_order)
% Smallest decay rate per component (:175-179, :256-259).
decay = ones(1,size(gt,1))*min(sA{1}(:));
keys = unique([gt(2:end,:) mom_order(2:end)'],'rows');
for j = 2:size(gt,1)
    q = find(ismember(keys,[gt(j,:) mom_order(j)],'rows'));
    decay(j) = min(sT{q}(:));
end
end

function [dt_b, la0] = stable_dt(par, M, even, odd, decay, dt, tleft)
% Bracketing and geometric bisection of the largest stable step (:180-203).
la = growth_factor(par,M,even,odd,decay,dt);
dt_b = [1 1]*dt; la_b = [1 1]*la; la0 = 1+1e-8;
while ~diff(la_b<la0) && (dt_b(1)<=tleft || la_b(1)>la0)
    ind = (la_b(1)<la0)+1;
    dt_b = dt_b(ind)*[1 2]*2^(ind-2);
    la_b = la_b([2 1]);
    la_b(ind) = growth_factor(par,M,even,odd,decay,dt_b(ind));
    if min(la_b)<1, la0 = 1; end
end
while dt_b(1)/dt_b(2)<0.98 && dt_b(1)<=tleft
    dt_mid = sqrt(prod(dt_b));
    la_mid = growth_factor(par,M,even,odd,decay,dt_mid);
    ind = 1+(la_mid>=la0);
    dt_b(ind) = dt_mid; la_b(ind) = la_mid;
end
dt_b = dt_b(1);
end

function la = growth_factor(par, M, even, odd, decay, dt)
% Largest eigenvalue modulus of the step's amplification matrix (:506-514).
G = diag(dt*expm1div(-decay*dt/2))*(M+diag(-decay/2));
Ge = G; Ge(odd,:) = 0; Ge = Ge+eye(size(G));
Go = G; Go(even,:) = 0; Go = Go+eye(size(G));
la = max(abs(eig(full(Go*Ge*Ge*Go))));
if par.timestep==2, la = la*exp(decay(1)*dt); end
end

function par = call_output(par, Uplot, gt, g, step)
% The output routine with the reference's argument lists (:429-472).
if ~isfield(par,'output') || isempty(par.output), return, end
comps = par.mom_output;
ys = g(2,gt(comps,2)); zs = g(3,gt(comps,3)); xs = g(1,gt(comps,1));
switch par.dim
    case 1
        for j = 1:numel(Uplot), Uplot{j} = reshape(Uplot{j},length(zs{j}),[]); end
        args = {zs, Uplot};
    case 2
        for j = 1:numel(Uplot), Uplot{j} = reshape(Uplot{j},length(ys{j}),[]); end
        args = {ys, zs, Uplot};
    otherwise
        args = {xs, ys, zs, Uplot};
end
if numel(comps)==1, args = cellfun(@(c)c{1},args,'Un',0); end
if nargout(par.output)
    par = par.output(par,args{:},step);
else
    par.output(par,args{:},step);
end
end

function y = expm1div(x)
% (exp(x)-1)/x with the series below |x| = 2e-4 (:531-535).
y = (exp(x)-1)./x;
s = abs(x)<=2e-4;
y(s) = 1+x(s)*.5+x(s).^2/6;
end

function f = capargs(fct, varargin)
% Call fct with as many leading arguments as it takes (:521-525).
f = fct(varargin{1:max(nargin(fct),1)});
end

function f = zero(varargin)
f = zeros(size(varargin{1}));
end
